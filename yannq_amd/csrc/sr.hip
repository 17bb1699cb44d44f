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
#include <algorithm>

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

// partial[k][c] = sum over samples of chunk k of w_s X[s][c] (w_s = 1 when w
// is null; X[s][c]^2 instead of X[s][c] when square), in sample order
__global__ void k_colsum(const double* __restrict__ X, int ns, int dim,
                         const double* __restrict__ w, int square,
                         double* __restrict__ partial) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= dim) return;
  const int k = blockIdx.y;
  const int s0 = k * kChunk, s1 = min(ns, s0 + kChunk);
  double acc = 0.0;
  for (int s = s0; s < s1; ++s) {
    const double x = X[(size_t)s * dim + c];
    const double v = square ? x * x : x;
    acc += w ? w[s] * v : v;
  }
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

// O_c = rs_s (O - omean) (rs_s = 1 when rs is null) and partial sums of
// rs_s O_c e_s
__global__ void k_center_force(double* __restrict__ O,
                               const double* __restrict__ omean,
                               const double* __restrict__ e, int ns, int dim,
                               const double* __restrict__ rs,
                               double* __restrict__ partial) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= dim) return;
  const int k = blockIdx.y;
  const int s0 = k * kChunk, s1 = min(ns, s0 + kChunk);
  const double mu = omean[c];
  double acc = 0.0;
  for (int s = s0; s < s1; ++s) {
    const double f = rs ? rs[s] : 1.0;
    const double v = (O[(size_t)s * dim + c] - mu) * f;
    O[(size_t)s * dim + c] = v;
    acc += v * (f * e[s]);
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

// ObsAvg<Matrix> (Runners/RunRBM.hpp:13-36) on the lower triangle of a
// column-major dim x dim S (nc doubles per entry): avg = beta avg +
// (1 - beta) S, then S = avg / corr with corr = 1 - beta^t (getAvg)
__global__ void k_ema_lower(double* __restrict__ avg, double* __restrict__ S, size_t dim, int nc,
                            double beta, double corr) {
  const size_t total = dim * dim;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (size_t)gridDim.x * blockDim.x) {
    if (q % dim < q / dim) continue;
    for (int k = 0; k < nc; ++k) {
      const size_t e = q * nc + k;
      const double a = beta * avg[e] + (1.0 - beta) * S[e];
      avg[e] = a;
      S[e] = a / corr;
    }
  }
}

// M = float(S) + shift I on the lower triangle (column-major dim x dim)
__global__ void k_lower_shift_f(const double* __restrict__ S, float* __restrict__ M, size_t dim,
                                double shift) {
  const size_t total = dim * dim;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (size_t)gridDim.x * blockDim.x) {
    const size_t r = q % dim, c = q / dim;
    if (r >= c) M[q] = (float)(S[q] + (r == c ? shift : 0.0));
  }
}

// S = sum_p P[p] over the lower triangle (column-major dim x dim, nc doubles
// per entry), partial products in index order
__global__ void k_sum_lower(const double* __restrict__ P, int np, size_t dim, int nc,
                            double* __restrict__ S) {
  const size_t total = dim * dim;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (size_t)gridDim.x * blockDim.x) {
    const size_t r = q % dim, c = q / dim;
    if (r < c) continue;
    for (int k = 0; k < nc; ++k) {
      double acc = 0.0;
      for (int p = 0; p < np; ++p) acc += P[((size_t)p * total + q) * nc + k];
      S[q * nc + k] = acc;
    }
  }
}

__global__ void k_add_diag(double* M, int dim, double shift) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < dim) M[(size_t)i * dim + i] += shift;
}

// ---- factored products on FP64 MFMA ---------------------------------------
// v_mfma_f64_16x16x4f64: lane l holds A[row l%16][k l/16] and B[k l/16][col
// l%16] and returns D[(l/16) + 4*q][l%16], q = 0..3 (rows interleaved across
// the lane groups, unlike the f32-output forms).  The augmented operands are
// produced on the fly from the int8 spins, T and the parameter vector.
typedef double doublex4 __attribute__((ext_vector_type(4)));
constexpr int kF64Tiles = 4;                     // column tiles per forward group
constexpr int kMaxIt64 = 9;                      // row tiles per backward wave

__device__ inline double sig_aug(const int8_t* __restrict__ sigma, int ns, int n, int s, int i) {
  if (s >= ns || i > n) return 0.0;
  if (i == n) return 1.0;
  return sigma[(size_t)s * n + i] > 0 ? 1.0 : -1.0;
}
// T' = [T | tb]: tb = 1 for a real machine (and the real plane of a complex
// one), 0 for the imaginary plane of a complex machine's T
__device__ inline double t_aug(const double* __restrict__ T, int ns, int m, int s, int j,
                               double tb = 1.0) {
  if (s >= ns || j > m) return 0.0;
  if (j == m) return tb;
  return T[(size_t)s * m + j];
}
__device__ inline double v_aug(const double* __restrict__ v, int n, int m, int use_bias, int k,
                               int j) {
  if (k < n && j < m) return v[(size_t)k * m + j];
  if (!use_bias) return 0.0;
  if (k < n && j == m) return v[(size_t)n * m + k];
  if (k == n && j < m) return v[(size_t)n * m + n + j];
  return 0.0;
}

// the parameter vector on the padded (Kp x Np) grid of sigma'^T T', so the
// forward product reads its B operand with plain coalesced loads
__global__ void k_aug64(const double* __restrict__ v, int n, int m, int use_bias, int Kp, int Np,
                        double* __restrict__ va) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)Kp * Np) return;
  va[q] = v_aug(v, n, m, use_bias, (int)(q / Np), (int)(q % Np));
}

// ypart[cg][s] = sum over the column tiles of group cg of T'_sj (sigma' V')_sj.
// Block: 4 waves x two 16-sample row tiles (128 samples), column group
// blockIdx.y of kF64Tiles tiles.  The group's rows of Vaug (the padded
// parameter grid) are staged through LDS kF64Kc k at a time and shared by the
// waves, and every B operand read from LDS feeds the MFMAs of both row tiles;
// the A operands are the lane's spins (+-1, 1 for the bias column, 0 past
// it).  Each (row tile, column tile) sum over k runs in increasing k and the
// column tiles are added in order, so the values are those of a sequential
// loop over the tiles.
constexpr int kF64Kc = 32;  // k per LDS stage (8 k-steps)
__global__ void __launch_bounds__(256)
    k_fwd64_mfma(const int8_t* __restrict__ sigma, const double* __restrict__ T, int ns, int n,
                 int m, const double* __restrict__ va, int Kp, int Np, int nrows,
                 double* __restrict__ ypart, double tb) {
  constexpr int kCols = kF64Tiles * 16;
  __shared__ double vsh[kF64Kc][kCols + 1];
  const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
  const int li = l & 15, lg = l >> 4;
  const int row0 = (blockIdx.x * 8 + wave * 2) * 16;  // this wave's two row tiles
  const int ct0 = blockIdx.y * kF64Tiles;
  const int nct = min(kF64Tiles, Np / 16 - ct0);
  const int nk = Kp / 4;
  // (waves past nrows run on clamped rows, for the barriers, and store nothing)
  bool rin[2];
  const int8_t* sp[2];
#pragma unroll
  for (int rt = 0; rt < 2; ++rt) {
    const int srow = row0 + rt * 16 + li;
    rin[rt] = srow < ns;
    sp[rt] = sigma + (size_t)(rin[rt] ? srow : 0) * n;
  }
  doublex4 c[2][kF64Tiles];
#pragma unroll
  for (int rt = 0; rt < 2; ++rt)
#pragma unroll
    for (int q = 0; q < kF64Tiles; ++q) c[rt][q] = (doublex4){0.0, 0.0, 0.0, 0.0};
  // the next stage's slab entries and spins are loaded into registers while
  // the current stage's MFMAs run
  constexpr int kPer = kF64Kc * kCols / 256;
  double stg[kPer];
  int8_t an[2][kF64Kc / 4];
  auto load = [&](int k0) {
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int q = tid + 256 * u, kk = q / kCols, jj = q % kCols;
      const int k = min(k0 + kk, Kp - 1), j = ct0 * 16 + min(jj, nct * 16 - 1);
      stg[u] = va[(size_t)k * Np + j];  // clamped: entries past Kp / the group are not read
    }
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int ks = 0; ks < kF64Kc / 4; ++ks) an[rt][ks] = sp[rt][min(k0 + 4 * ks + lg, n - 1)];
  };
  load(0);
  for (int k0 = 0; k0 < Kp; k0 += kF64Kc) {
    __syncthreads();  // the previous stage is read
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int q = tid + 256 * u;
      vsh[q / kCols][q % kCols] = stg[u];
    }
    int8_t av[2][kF64Kc / 4];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int ks = 0; ks < kF64Kc / 4; ++ks) av[rt][ks] = an[rt][ks];
    __syncthreads();
    if (k0 + kF64Kc < Kp) load(k0 + kF64Kc);
#pragma unroll
    for (int ks = 0; ks < kF64Kc / 4; ++ks) {
      if (k0 / 4 + ks >= nk) break;  // uniform
      const int i = k0 + 4 * ks + lg;  // this lane's spin index
      double a[2];
#pragma unroll
      for (int rt = 0; rt < 2; ++rt)
        a[rt] = (!rin[rt] || i > n) ? 0.0 : (i == n ? 1.0 : (av[rt][ks] > 0 ? 1.0 : -1.0));
#pragma unroll
      for (int q = 0; q < kF64Tiles; ++q) {
        if (q >= nct) break;  // uniform
        const double b = vsh[4 * ks + lg][q * 16 + li];
#pragma unroll
        for (int rt = 0; rt < 2; ++rt)
          c[rt][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt], b, c[rt][q], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int rt = 0; rt < 2; ++rt) {
    const int r0 = row0 + rt * 16;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < kF64Tiles; ++q)
      if (q < nct) {
        const int j = (ct0 + q) * 16 + li;
#pragma unroll
        for (int e = 0; e < 4; ++e)
          acc[e] += t_aug(T, ns, m, r0 + lg + 4 * e, j, tb) * c[rt][q][e];
      }
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
      for (int s = 8; s >= 1; s >>= 1) acc[e] += __shfl_xor(acc[e], s);
    if (li == 0 && r0 < nrows)
#pragma unroll
      for (int e = 0; e < 4; ++e) ypart[(size_t)blockIdx.y * nrows + r0 + lg + 4 * e] = acc[e];
  }
}

__global__ void k_sum_ypart(const double* __restrict__ ypart, int ncg, int nrows, int ns,
                            double* __restrict__ y) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  double v = 0.0;
  for (int g = 0; g < ncg; ++g) v += ypart[(size_t)g * nrows + s];
  y[s] = v;
}

constexpr int kU64 = 8;  // k-steps per load batch of the backward product

// part[kc][i][j] = sum_{s in chunk kc} sigma'_si r_s T'_sj.  One wave per
// (column tile, sample chunk blockIdx.y, group of NIT row tiles blockIdx.z).
// Per k-step a lane needs its B value r_s T'_sj and the spins of its NIT rows
// of sample s; both are loaded a batch of kU64 k-steps ahead of the MFMAs
// that use them.  NIT is a template parameter so that the row-tile loop
// unrolls and the accumulators stay in registers.  Each tile's sum over the
// samples runs in increasing sample order.
template <int NIT>
__global__ void __launch_bounds__(256)
    k_bwd64_mfma(const int8_t* __restrict__ sigma, const double* __restrict__ T, int ns, int n,
                 int m, const double* __restrict__ r, int Kp, int Np, int chunk,
                 double* __restrict__ part, double tb) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int ct = blockIdx.x * 4 + wave;
  if (ct >= Np / 16) return;  // wave-uniform; no barrier in this kernel
  const int li = l & 15, lg = l >> 4;
  const int kc = blockIdx.y;
  const int i0 = blockIdx.z * NIT * 16;  // first row of G' formed here
  const int s0 = kc * chunk;
  const int j = ct * 16 + li;
  doublex4 acc[NIT];
#pragma unroll
  for (int it = 0; it < NIT; ++it) acc[it] = (doublex4){0.0, 0.0, 0.0, 0.0};
  const int nks = chunk / 4;
  // spin (row i0 + it * 16 + li, sample s) as +-1, 1 for the bias row, 0 past
  // it or past the batch: loaded at clamped addresses, selected afterwards
  double bv[kU64];
  int8_t av[kU64][NIT];
  auto load = [&](int k0, double (&ob)[kU64], int8_t (&oa)[kU64][NIT]) {
#pragma unroll
    for (int u = 0; u < kU64; ++u) {
      const int s = s0 + min(k0 + u, nks - 1) * 4 + lg;  // clamped: unused past nks
      const int sc = min(s, ns - 1);
      ob[u] = (s < ns ? r[s] : 0.0) * t_aug(T, ns, m, s, j, tb);
#pragma unroll
      for (int it = 0; it < NIT; ++it)
        oa[u][it] = sigma[(size_t)sc * n + min(i0 + it * 16 + li, n - 1)];
    }
  };
  auto spin = [&](int8_t raw, int it, int k) -> double {
    const int i = i0 + it * 16 + li;
    const int s = s0 + k * 4 + lg;
    if (s >= ns || i > n) return 0.0;
    if (i == n) return 1.0;
    return raw > 0 ? 1.0 : -1.0;
  };
  load(0, bv, av);
  for (int k0 = 0; k0 < nks; k0 += kU64) {
    double bn[kU64];
    int8_t an[kU64][NIT];
    if (k0 + kU64 < nks) load(k0 + kU64, bn, an);
#pragma unroll
    for (int u = 0; u < kU64; ++u) {
      if (k0 + u >= nks) break;  // uniform
#pragma unroll
      for (int it = 0; it < NIT; ++it)
        acc[it] = __builtin_amdgcn_mfma_f64_16x16x4f64(spin(av[u][it], it, k0 + u), bv[u],
                                                       acc[it], 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < kU64; ++u) {
      bv[u] = bn[u];
#pragma unroll
      for (int it = 0; it < NIT; ++it) av[u][it] = an[u][it];
    }
  }
  double* out = part + (size_t)kc * Kp * Np;
#pragma unroll
  for (int it = 0; it < NIT; ++it)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int row = i0 + it * 16 + lg + 4 * q;
      if (row < Kp) out[(size_t)row * Np + j] = acc[it][q];
    }
}

// g[c] = scale * sum_k part[k][aug(c)] on the padded (Kp x Np) grid, mapped to
// the [W | a | b] layout
__global__ void k_reduce_aug(const double* __restrict__ part, int chunks, int n, int m,
                             int use_bias, int Kp, int Np, double scale,
                             double* __restrict__ g) {
  const size_t nm = (size_t)n * m;
  const size_t dim = use_bias ? nm + n + m : nm;
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= dim) return;
  size_t a;
  if (c < nm) a = (c / m) * Np + (c % m);
  else if (c < nm + n) a = (c - nm) * Np + m;
  else a = (size_t)n * Np + (c - nm - n);
  const size_t stride = (size_t)Kp * Np;
  // the chunks in order, 16 loads in flight at a time (clamped indices, the
  // values past the last chunk left out of the sum)
  constexpr int kB = 16;
  double acc = 0.0;
  for (int k0 = 0; k0 < chunks; k0 += kB) {
    double v[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) v[u] = part[(size_t)min(k0 + u, chunks - 1) * stride + a];
#pragma unroll
    for (int u = 0; u < kB; ++u)
      if (k0 + u < chunks) acc += v[u];
  }
  g[c] = acc * scale;
}

inline int kp16(int n) { return (n + 1 + 15) / 16 * 16; }
inline int np16(int m) { return (m + 1 + 15) / 16 * 16; }

// partial[k][j] = sum over the samples of chunk k of T[s][j]^2
__global__ void k_colsum_sq(const double* __restrict__ T, int ns, int m,
                            double* __restrict__ partial) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const int k = blockIdx.y;
  const int s0 = k * kChunk, s1 = min(ns, s0 + kChunk);
  double acc = 0.0;
  for (int s = s0; s < s1; ++s) {
    const double t = T[(size_t)s * m + j];
    acc += t * t;
  }
  partial[(size_t)k * m + j] = acc;
}

// diag(S) of the factored O: Var(O_c) with O_W(i,j)^2 = T_j^2 and O_a^2 = 1
__global__ void k_diag_S(const double* __restrict__ meanT2, const double* __restrict__ omean,
                         int n, int m, int use_bias, double* __restrict__ diag) {
  const size_t nm = (size_t)n * m;
  const size_t dim = use_bias ? nm + n + m : nm;
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= dim) return;
  const double mu = omean[c];
  double e2;
  if (c < nm) e2 = meanT2[c % m];
  else if (c < nm + n) e2 = 1.0;
  else e2 = meanT2[c - nm - n];
  diag[c] = e2 - mu * mu;
}

__global__ void k_jacobi_inv64(const double* __restrict__ diag, double shift,
                               double* __restrict__ out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = 1.0 / (diag[i] + shift);
}

__global__ void k_mul(const double* __restrict__ a, const double* __restrict__ b,
                      double* __restrict__ out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] * b[i];
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
// Device-resident scalars for operators that run without host round trips:
// out = the sum of nb dot partials in block order (staged in LDS, added by one
// thread: the order of the host loop that reads them back), then the same
// updates as k_add_const and k_axpy with the scalar (times sign) read on the
// device.
__global__ void __launch_bounds__(256)
    k_psum_scalar(const double* __restrict__ partial, int nb, double* __restrict__ out) {
  __shared__ double sh[512];
  double s = 0.0;
  for (int base = 0; base < nb; base += 512) {
    const int cnt = min(512, nb - base);
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) sh[i] = partial[base + i];
    __syncthreads();
    if (threadIdx.x == 0)
      for (int i = 0; i < cnt; ++i) s += sh[i];
  }
  if (threadIdx.x == 0) *out = s;
}
__global__ void k_add_const_p(double* y, const double* __restrict__ c, double sign, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += sign * *c;
}
__global__ void k_axpy_p(const double* __restrict__ a, double sign,
                         const double* __restrict__ x, double* __restrict__ y, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += (sign * *a) * x[i];
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
// samples per chunk of the FP64 backward product: 64 for batches of up to
// 4096 samples (four times the waves of 256-sample chunks; each wave's k-steps
// are a chain of dependent MFMAs), else 256
int bwd64_chunk(int ns) { return ns <= 4096 ? 64 : kChunk; }
int bwd64_chunks(int ns) { return (ns + bwd64_chunk(ns) - 1) / bwd64_chunk(ns); }

void launch_colsum(const double* X, int ns, int dim, double* partial,
                   hipStream_t st, const double* w, int square) {
  hipLaunchKernelGGL(k_colsum, dim3(nblk(dim), colsum_chunks(ns)), dim3(256), 0, st,
                     X, ns, dim, w, square, partial);
}

void launch_reduce_partials(const double* partial, int chunks, int dim,
                            double scale, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_partials, dim3(nblk(dim)), dim3(256), 0, st, partial,
                     chunks, dim, scale, out);
}

void launch_center_force(double* O, const double* omean, const double* e,
                         int ns, int dim, double* partial, hipStream_t st, const double* rs) {
  hipLaunchKernelGGL(k_center_force, dim3(nblk(dim), colsum_chunks(ns)), dim3(256),
                     0, st, O, omean, e, ns, dim, rs, partial);
}

void launch_sum_lower(const double* P, int np, size_t dim, int nc, double* S, hipStream_t st) {
  const size_t total = dim * dim;
  const unsigned grid = (unsigned)std::min<size_t>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(k_sum_lower, dim3(grid), dim3(256), 0, st, P, np, dim, nc, S);
}

void launch_ema_lower(double* avg, double* S, size_t dim, int nc, double beta, double corr,
                      hipStream_t st) {
  const size_t total = dim * dim;
  const unsigned grid = (unsigned)std::min<size_t>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(k_ema_lower, dim3(grid), dim3(256), 0, st, avg, S, dim, nc, beta, corr);
}

void launch_lower_shift_f(const double* S, float* M, size_t dim, double shift, hipStream_t st) {
  const size_t total = dim * dim;
  const unsigned grid = (unsigned)std::min<size_t>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(k_lower_shift_f, dim3(grid), dim3(256), 0, st, S, M, dim, shift);
}

void launch_add_diag(double* M, int dim, double shift, hipStream_t st) {
  hipLaunchKernelGGL(k_add_diag, dim3(nblk(dim)), dim3(256), 0, st, M, dim, shift);
}

void launch_copy_lower_shift(const double* S, double* M, int dim,
                             double shift, hipStream_t st) {
  hipLaunchKernelGGL(k_copy_lower_shift, dim3(8192), dim3(256), 0, st, S, M, dim,
                     shift);
}

size_t factored_partial_doubles(int ns, int n, int m) {
  const size_t grid = (size_t)bwd64_chunks(ns) * kp16(n) * np16(m);
  // forward: the column-group partials, then the padded parameter grid
  const size_t fwd = (size_t)((np16(m) / 16 + kF64Tiles - 1) / kF64Tiles) * ((ns + 15) / 16 * 16) +
                     (size_t)kp16(n) * np16(m);
  return (grid > fwd ? grid : fwd) + ns;
}

void launch_factored_fwd64(const int8_t* sigma, const double* T, int ns, int n,
                           int m, int use_bias, const double* v, double* y,
                           double* scratch, hipStream_t st, double tb) {
  const int Kp = kp16(n), Np = np16(m), nrows = (ns + 15) / 16 * 16;
  const int ncg = (Np / 16 + kF64Tiles - 1) / kF64Tiles;
  double* va = scratch + (size_t)ncg * nrows;
  ProfGuard g("factored_fwd64", st);
  hipLaunchKernelGGL(k_aug64, dim3(nblk((size_t)Kp * Np)), dim3(256), 0, st, v, n, m, use_bias,
                     Kp, Np, va);
  hipLaunchKernelGGL(k_fwd64_mfma, dim3((nrows + 127) / 128, ncg), dim3(256), 0, st, sigma, T,
                     ns, n, m, (const double*)va, Kp, Np, nrows, scratch, tb);
  hipLaunchKernelGGL(k_sum_ypart, dim3(nblk(ns)), dim3(256), 0, st, scratch, ncg, nrows, ns, y);
}

// the backward kernel for nit row tiles per block
auto bwd64_kernel(int nit) {
  switch (nit) {
    case 1: return k_bwd64_mfma<1>;
    case 2: return k_bwd64_mfma<2>;
    case 3: return k_bwd64_mfma<3>;
    case 4: return k_bwd64_mfma<4>;
    case 5: return k_bwd64_mfma<5>;
    case 6: return k_bwd64_mfma<6>;
    case 7: return k_bwd64_mfma<7>;
    case 8: return k_bwd64_mfma<8>;
    default: return k_bwd64_mfma<kMaxIt64>;
  }
}

void launch_factored_bwd64(const int8_t* sigma, const double* T, int ns, int n,
                           int m, int use_bias, const double* r, double* g,
                           double* partial, hipStream_t st, double tb) {
  const int chunks = bwd64_chunks(ns);
  const int Kp = kp16(n), Np = np16(m);
  ProfGuard pg("factored_bwd64", st);
  // all row tiles in one group up to kMaxIt64, else groups of kMaxIt64
  const int nit = std::min(Kp / 16, kMaxIt64);
  const int rgroups = (Kp / 16 + nit - 1) / nit;
  hipLaunchKernelGGL(bwd64_kernel(nit), dim3((Np / 16 + 3) / 4, chunks, rgroups), dim3(256), 0,
                     st, sigma, T, ns, n, m, r, Kp, Np, bwd64_chunk(ns), partial, tb);
  const size_t dim = use_bias ? (size_t)n * m + n + m : (size_t)n * m;
  hipLaunchKernelGGL(k_reduce_aug, dim3(nblk(dim)), dim3(256), 0, st, partial, chunks, n, m,
                     use_bias, Kp, Np, 1.0, g);
}

void launch_factored_mean64(const int8_t* sigma, const double* T, int ns,
                            int n, int m, int use_bias, double* omean,
                            double* partial, hipStream_t st, double tb) {
  // r = 1/ns lives in the tail of the partial buffer
  double* ones = partial + factored_partial_doubles(ns, n, m) - ns;
  hipLaunchKernelGGL(k_fill, dim3(nblk(ns)), dim3(256), 0, st, ones, 1.0 / ns, ns);
  launch_factored_bwd64(sigma, T, ns, n, m, use_bias, ones, omean, partial, st, tb);
}

void launch_diag_S(const double* T, int ns, int n, int m, int use_bias, const double* omean,
                   double* diag, double* partial, hipStream_t st) {
  const int chunks = colsum_chunks(ns);
  hipLaunchKernelGGL(k_colsum_sq, dim3(nblk(m), chunks), dim3(256), 0, st, T, ns, m, partial);
  double* meanT2 = partial + (size_t)chunks * m;
  hipLaunchKernelGGL(k_reduce_partials, dim3(nblk(m)), dim3(256), 0, st, partial, chunks, m,
                     1.0 / ns, meanT2);
  const size_t dim = use_bias ? (size_t)n * m + n + m : (size_t)n * m;
  hipLaunchKernelGGL(k_diag_S, dim3(nblk(dim)), dim3(256), 0, st, meanT2, omean, n, m, use_bias,
                     diag);
}

void launch_jacobi_inv64(const double* diag, double shift, double* out, size_t n,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_jacobi_inv64, dim3(nblk(n)), dim3(256), 0, st, diag, shift, out, n);
}

void launch_mul(const double* a, const double* b, double* out, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_mul, dim3(nblk(n)), dim3(256), 0, st, a, b, out, n);
}

int dot_blocks(size_t n) {
  const size_t b = (n + 255) / 256;
  return (int)(b < 512 ? (b ? b : 1) : 512);
}

void launch_dot(const double* x, const double* y, size_t n, double* partial,
                hipStream_t st) {
  hipLaunchKernelGGL(k_dot, dim3(dot_blocks(n)), dim3(256), 0, st, x, y, n, partial);
}
void launch_dot_dev(const double* x, const double* y, size_t n, double* partial, double* out,
                    hipStream_t st) {
  launch_dot(x, y, n, partial, st);
  hipLaunchKernelGGL(k_psum_scalar, dim3(1), dim3(256), 0, st, partial, dot_blocks(n), out);
}
void launch_add_const_p(double* y, const double* c, double sign, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_add_const_p, dim3(nblk(n)), dim3(256), 0, st, y, c, sign, n);
}
void launch_axpy_p(const double* a, double sign, const double* x, double* y, size_t n,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_axpy_p, dim3(nblk(n)), dim3(256), 0, st, a, sign, x, y, n);
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
