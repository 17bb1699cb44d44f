// SR building blocks in FP64: O rows (GroundState/utils.hpp:13-51 with
// RBM::logDeriv, RBM.hpp:390-406), column means, centring and the force
// F = O_c^T E_loc / Ns (SRMat.hpp:89-103), and the factored products O v and
// O^T r that never materialise O.
//
// Factored form.  With sigma' = [sigma | 1] (n+1 columns) and
// T' = [tanh(theta) | 1] (m+1 columns), every row of O is the Kronecker
// product sigma'_s (x) T'_s without its (n, m) corner: the W block
// sigma_i T_j, the a block sigma_i (j = m), the b block T_j (i = n).  Hence
//   (O v)_s   = sum_{i,j} sigma'_si V'_ij T'_sj,
//   (O^T r)_ij = sum_s sigma'_si r_s T'_sj,
// with V' the parameter vector laid out on the (n+1) x (m+1) grid.  All
// reductions here use fixed orders, so results are run-to-run identical.
#include "yq_internal.h"

namespace yq {
namespace {

constexpr int kChunk = 256;  // samples per partial-sum chunk

template <typename R>
__global__ void k_tanh(const R* __restrict__ th, R* __restrict__ T, size_t cnt) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) T[i] = tanh(th[i]);
}

__global__ void k_build_O(const int8_t* __restrict__ sigma,
                          const double* __restrict__ T, int ns, int n, int m,
                          int use_bias, double* __restrict__ O) {
  const size_t nm = (size_t)n * m;
  const size_t dim = use_bias ? nm + n + m : nm;
  const size_t total = (size_t)ns * dim;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (size_t)gridDim.x * blockDim.x) {
    const size_t s = q / dim, c = q % dim;
    double v;
    if (c < nm) {
      const int i = (int)(c / m), j = (int)(c % m);
      v = (sigma[s * n + i] > 0 ? 1.0 : -1.0) * T[s * m + j];
    } else if (c < nm + n) {
      v = sigma[s * n + (c - nm)] > 0 ? 1.0 : -1.0;
    } else {
      v = T[s * m + (c - nm - n)];
    }
    O[q] = v;
  }
}

// partial[k][c] = sum over samples of chunk k of X[s][c], in sample order
__global__ void k_colsum(const double* __restrict__ X, int ns, int dim,
                         double* __restrict__ partial) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= dim) return;
  const int k = blockIdx.y;
  const int s0 = k * kChunk, s1 = min(ns, s0 + kChunk);
  double acc = 0.0;
  for (int s = s0; s < s1; ++s) acc += X[(size_t)s * dim + c];
  partial[(size_t)k * dim + c] = acc;
}

__global__ void k_reduce_partials(const double* __restrict__ partial,
                                  int chunks, int dim, double scale,
                                  double* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= dim) return;
  double acc = 0.0;
  for (int k = 0; k < chunks; ++k) acc += partial[(size_t)k * dim + c];
  out[c] = acc * scale;
}

__global__ void k_center_force(double* __restrict__ O,
                               const double* __restrict__ omean,
                               const double* __restrict__ e, int ns, int dim,
                               double* __restrict__ partial) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= dim) return;
  const int k = blockIdx.y;
  const int s0 = k * kChunk, s1 = min(ns, s0 + kChunk);
  const double mu = omean[c];
  double acc = 0.0;
  for (int s = s0; s < s1; ++s) {
    const double v = O[(size_t)s * dim + c] - mu;
    O[(size_t)s * dim + c] = v;
    acc += v * e[s];
  }
  partial[(size_t)k * dim + c] = acc;
}

__global__ void k_copy_lower_shift(const double* __restrict__ S,
                                   double* __restrict__ M, int dim,
                                   double shift) {
  const size_t total = (size_t)dim * dim;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (size_t)gridDim.x * blockDim.x) {
    const size_t r = q % dim, c = q / dim;  // column-major
    M[q] = S[q] + (r == c ? shift : 0.0);
  }
}

__global__ void k_add_diag(double* M, int dim, double shift) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < dim) M[(size_t)i * dim + i] += shift;
}

// ---- factored products --------------------------------------------------
constexpr int kFwdS = 8;  // samples per block in the forward product

// y_s = sum_{i,j} sigma'_si V'_ij T'_sj (uncentred), V' read from v in the
// [W | a | b] layout.
__global__ void __launch_bounds__(256)
    k_factored_fwd64(const int8_t* __restrict__ sigma,
                     const double* __restrict__ T, int ns, int n, int m,
                     int use_bias, const double* __restrict__ v,
                     double* __restrict__ y) {
  __shared__ int8_t sg[kFwdS][kMaxSites];
  __shared__ double red[kFwdS][256];
  const int s0 = blockIdx.x * kFwdS;
  const int cnt = min(kFwdS, ns - s0);
  for (int q = threadIdx.x; q < cnt * n; q += 256)
    sg[q / n][q % n] = sigma[(size_t)s0 * n + q];
  __syncthreads();
  const size_t nm = (size_t)n * m;
  double acc[kFwdS];
#pragma unroll
  for (int t = 0; t < kFwdS; ++t) acc[t] = 0.0;
  // columns j < m: u_sj = sum_i sigma_si V_ij (+ b_j), then times T_sj
  for (int j = threadIdx.x; j < m; j += 256) {
    double u[kFwdS];
    const double bj = use_bias ? v[nm + n + j] : 0.0;
#pragma unroll
    for (int t = 0; t < kFwdS; ++t) u[t] = 0.0;
    for (int i = 0; i < n; ++i) {
      const double w = v[(size_t)i * m + j];
#pragma unroll
      for (int t = 0; t < kFwdS; ++t) u[t] = sg[t][i] > 0 ? u[t] + w : u[t] - w;
    }
#pragma unroll
    for (int t = 0; t < kFwdS; ++t)
      if (t < cnt) acc[t] += T[(size_t)(s0 + t) * m + j] * (u[t] + bj);
  }
  // column j = m (the a block): sum_i sigma_si a_i
  if (use_bias)
    for (int i = threadIdx.x; i < n; i += 256) {
      const double w = v[nm + i];
#pragma unroll
      for (int t = 0; t < kFwdS; ++t) acc[t] += sg[t][i] > 0 ? w : -w;
    }
#pragma unroll
  for (int t = 0; t < kFwdS; ++t) red[t][threadIdx.x] = acc[t];
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int t = 0; t < kFwdS; ++t) red[t][threadIdx.x] += red[t][threadIdx.x + w];
    __syncthreads();
  }
  if ((int)threadIdx.x < cnt) y[s0 + threadIdx.x] = red[threadIdx.x][0];
}

// partial[k][i*(m+1)+j] = sum_{s in chunk k} sigma'_si r_s T'_sj on the
// augmented grid.  Block: 64 columns j x 4 row groups; each thread owns rows
// i = g, g+4, ... (at most kMaxSites/4 + 1 accumulators).
constexpr int kBwdRows = kMaxSites / 4 + 1;
__global__ void __launch_bounds__(256)
    k_factored_bwd64(const int8_t* __restrict__ sigma,
                     const double* __restrict__ T, int ns, int n, int m,
                     const double* __restrict__ r,
                     double* __restrict__ partial) {
  __shared__ int8_t sg[kChunk][kMaxSites];
  __shared__ double rr[kChunk];
  const int k = blockIdx.y;
  const int s0 = k * kChunk, cnt = min(kChunk, ns - s0);
  for (int q = threadIdx.x; q < cnt * n; q += 256)
    sg[q / n][q % n] = sigma[(size_t)s0 * n + q];
  for (int q = threadIdx.x; q < cnt; q += 256) rr[q] = r[s0 + q];
  __syncthreads();
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int g = threadIdx.x >> 6;
  if (j > m) return;
  double acc[kBwdRows];
#pragma unroll
  for (int q = 0; q < kBwdRows; ++q) acc[q] = 0.0;
  for (int s = 0; s < cnt; ++s) {
    const double tq = (j < m ? T[(size_t)(s0 + s) * m + j] : 1.0) * rr[s];
#pragma unroll
    for (int q = 0; q < kBwdRows; ++q) {
      const int i = g + 4 * q;
      if (i < n) acc[q] += sg[s][i] > 0 ? tq : -tq;
      else if (i == n) acc[q] += tq;
    }
  }
  const size_t W1 = (size_t)(m + 1);
  double* out = partial + (size_t)k * (n + 1) * W1;
#pragma unroll
  for (int q = 0; q < kBwdRows; ++q) {
    const int i = g + 4 * q;
    if (i <= n) out[(size_t)i * W1 + j] = acc[q];
  }
}

// g[c] = scale * sum_k partial[k][aug(c)], mapping the augmented grid to the
// [W | a | b] layout
__global__ void k_reduce_aug(const double* __restrict__ partial, int chunks,
                             int n, int m, int use_bias, double scale,
                             double* __restrict__ g) {
  const size_t nm = (size_t)n * m;
  const size_t dim = use_bias ? nm + n + m : nm;
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= dim) return;
  size_t a;
  const size_t W1 = (size_t)(m + 1);
  if (c < nm) a = (c / m) * W1 + (c % m);
  else if (c < nm + n) a = (c - nm) * W1 + m;
  else a = (size_t)n * W1 + (c - nm - n);
  const size_t stride = (size_t)(n + 1) * W1;
  double acc = 0.0;
  for (int k = 0; k < chunks; ++k) acc += partial[(size_t)k * stride + a];
  g[c] = acc * scale;
}

__global__ void k_fill(double* x, double v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// ---- vectors --------------------------------------------------------------
__global__ void __launch_bounds__(256)
    k_dot(const double* __restrict__ x, const double* __restrict__ y, size_t n,
          double* __restrict__ partial) {
  __shared__ double red[256];
  double acc = 0.0;
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
    acc += x[i] * y[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_axpy(double a, const double* __restrict__ x,
                       double* __restrict__ y, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}
__global__ void k_xpby(const double* __restrict__ x, double b,
                       double* __restrict__ y, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i] + b * y[i];
}
__global__ void k_scale_add(double* __restrict__ y, double a,
                            const double* __restrict__ x, double b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * y[i] + b * x[i];
}
__global__ void k_add_const(double* y, double c, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += c;
}
__global__ void k_to_float(const double* __restrict__ x, float* __restrict__ y,
                           size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}
__global__ void k_to_double(const float* __restrict__ x, double* __restrict__ y,
                            size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (double)x[i];
}
template <typename O>
__global__ void k_transpose(const double* __restrict__ W, int m, int n,
                            O* __restrict__ WT) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)m * n) return;
  const size_t i = q / m, j = q % m;  // W[i*m + j] = W_ji
  WT[j * n + i] = (O)W[q];
}

inline unsigned nblk(size_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace

void launch_tanh(const void* theta, void* T, size_t count, int use_float,
                 hipStream_t st) {
  if (use_float)
    hipLaunchKernelGGL(k_tanh<float>, dim3(nblk(count)), dim3(256), 0, st,
                       static_cast<const float*>(theta), static_cast<float*>(T), count);
  else
    hipLaunchKernelGGL(k_tanh<double>, dim3(nblk(count)), dim3(256), 0, st,
                       static_cast<const double*>(theta), static_cast<double*>(T), count);
}

void launch_build_O(const int8_t* sigma, const double* T, int ns, int n, int m,
                    int use_bias, double* O, hipStream_t st) {
  ProfGuard g("build_O", st);
  hipLaunchKernelGGL(k_build_O, dim3(4096), dim3(256), 0, st, sigma, T, ns, n, m,
                     use_bias, O);
}

int colsum_chunks(int ns) { return (ns + kChunk - 1) / kChunk; }

void launch_colsum(const double* X, int ns, int dim, double* partial,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_colsum, dim3(nblk(dim), colsum_chunks(ns)), dim3(256), 0, st,
                     X, ns, dim, partial);
}

void launch_reduce_partials(const double* partial, int chunks, int dim,
                            double scale, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_partials, dim3(nblk(dim)), dim3(256), 0, st, partial,
                     chunks, dim, scale, out);
}

void launch_center_force(double* O, const double* omean, const double* e,
                         int ns, int dim, double* partial, hipStream_t st) {
  hipLaunchKernelGGL(k_center_force, dim3(nblk(dim), colsum_chunks(ns)), dim3(256),
                     0, st, O, omean, e, ns, dim, partial);
}

void launch_add_diag(double* M, int dim, double shift, hipStream_t st) {
  hipLaunchKernelGGL(k_add_diag, dim3(nblk(dim)), dim3(256), 0, st, M, dim, shift);
}

void launch_copy_lower_shift(const double* S, double* M, int dim,
                             double shift, hipStream_t st) {
  hipLaunchKernelGGL(k_copy_lower_shift, dim3(8192), dim3(256), 0, st, S, M, dim,
                     shift);
}

void launch_factored_fwd64(const int8_t* sigma, const double* T, int ns, int n,
                           int m, int use_bias, const double* v, double* y,
                           hipStream_t st) {
  ProfGuard g("factored_fwd64", st);
  hipLaunchKernelGGL(k_factored_fwd64, dim3((ns + kFwdS - 1) / kFwdS), dim3(256), 0,
                     st, sigma, T, ns, n, m, use_bias, v, y);
}

void launch_factored_bwd64(const int8_t* sigma, const double* T, int ns, int n,
                           int m, int use_bias, const double* r, double* g,
                           double* partial, hipStream_t st) {
  const int chunks = colsum_chunks(ns);
  ProfGuard pg("factored_bwd64", st);
  hipLaunchKernelGGL(k_factored_bwd64, dim3((m + 1 + 63) / 64, chunks), dim3(256), 0,
                     st, sigma, T, ns, n, m, r, partial);
  const size_t dim = use_bias ? (size_t)n * m + n + m : (size_t)n * m;
  hipLaunchKernelGGL(k_reduce_aug, dim3(nblk(dim)), dim3(256), 0, st, partial, chunks,
                     n, m, use_bias, 1.0, g);
}

void launch_factored_mean64(const int8_t* sigma, const double* T, int ns,
                            int n, int m, int use_bias, double* omean,
                            double* partial, hipStream_t st) {
  // r = 1/ns written into the tail of the partial buffer's first chunk is
  // avoided: use a dedicated fill of a scratch vector at the end of partial.
  const int chunks = colsum_chunks(ns);
  double* ones = partial + (size_t)chunks * (n + 1) * (m + 1);
  hipLaunchKernelGGL(k_fill, dim3(nblk(ns)), dim3(256), 0, st, ones, 1.0 / ns, ns);
  launch_factored_bwd64(sigma, T, ns, n, m, use_bias, ones, omean, partial, st);
}

int dot_blocks(size_t n) {
  const size_t b = (n + 255) / 256;
  return (int)(b < 512 ? (b ? b : 1) : 512);
}

void launch_dot(const double* x, const double* y, size_t n, double* partial,
                hipStream_t st) {
  hipLaunchKernelGGL(k_dot, dim3(dot_blocks(n)), dim3(256), 0, st, x, y, n, partial);
}
void launch_axpy(double alpha, const double* x, double* y, size_t n,
                 hipStream_t st) {
  hipLaunchKernelGGL(k_axpy, dim3(nblk(n)), dim3(256), 0, st, alpha, x, y, n);
}
void launch_xpby(const double* x, double beta, double* y, size_t n,
                 hipStream_t st) {
  hipLaunchKernelGGL(k_xpby, dim3(nblk(n)), dim3(256), 0, st, x, beta, y, n);
}
void launch_scale_add(double* y, double a, const double* x, double b, size_t n,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_scale_add, dim3(nblk(n)), dim3(256), 0, st, y, a, x, b, n);
}
void launch_add_const(double* y, double c, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_add_const, dim3(nblk(n)), dim3(256), 0, st, y, c, n);
}
void launch_to_float(const double* x, float* y, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_to_float, dim3(nblk(n)), dim3(256), 0, st, x, y, n);
}
void launch_float_to_double(const float* x, double* y, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_to_double, dim3(nblk(n)), dim3(256), 0, st, x, y, n);
}
void launch_transpose(const double* W, int m, int n, double* WT,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_transpose<double>, dim3(nblk((size_t)m * n)), dim3(256), 0, st,
                     W, m, n, WT);
}
void launch_transpose_f(const double* W, int m, int n, float* WT,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_transpose<float>, dim3(nblk((size_t)m * n)), dim3(256), 0, st,
                     W, m, n, WT);
}

}  // namespace yq
