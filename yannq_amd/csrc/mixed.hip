// Mixed-precision factored SR operator on gfx950 bf16 MFMA.
//
// O's rows are sigma'_s (x) T'_s (sigma' = [sigma | 1], T' = [tanh theta | 1],
// see sr.hip).  Spins are exactly representable in bf16, so every product
// keeps full float accuracy by splitting only the float operand into three
// bf16 pieces (x = hi + mid + lo to 24 bits) and issuing one MFMA per piece
// with float accumulation:
//   forward   U' = sigma' V'        (ns x Np, K = n+1 -> Kp)   then
//             y_s = sum_j T'_sj U'_sj           (fused epilogue)
//   backward  G' = sigma'^T diag(y) T'  (Kp x Np, K = ns, split in chunks)
// v_mfma_f32_16x16x32_bf16 operand layout used here: lane l holds
// A[row l%16][k 8*(l/16) .. +7], B[k 8*(l/16) .. +7][col l%16] and returns
// D[4*(l/16) + q][l%16], q = 0..3.
#include <math.h>

#include "mixed.h"
#include "yq_internal.h"

namespace yq {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

constexpr int kBwdChunk = 512;  // samples per backward chunk
constexpr int kMaxKs = 5;       // Kp <= 160 (n <= 128)
constexpr int kMaxIt = 10;      // Kp / 16

__device__ __host__ inline uint16_t f2bf(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __host__ inline float bf2f(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float x;
  memcpy(&x, &u, 4);
  return x;
}
__device__ inline void split3(float x, uint16_t& hi, uint16_t& mid, uint16_t& lo) {
  hi = f2bf(x);
  const float r1 = x - bf2f(hi);
  mid = f2bf(r1);
  const float r2 = r1 - bf2f(mid);
  lo = f2bf(r2);
}
__device__ inline bf16x8 ld8(const uint16_t* p) {
  const u16x8 v = *reinterpret_cast<const u16x8*>(p);
  return __builtin_bit_cast(bf16x8, v);
}

// sigma' (bf16) row-major and transposed; zero outside ns x (n+1)
__global__ void k_prep_sig(const int8_t* __restrict__ sigma, int ns, int n,
                           int Kp, int nsp, uint16_t* __restrict__ A,
                           uint16_t* __restrict__ AT) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)nsp * Kp) return;
  const int s = (int)(q / Kp), i = (int)(q % Kp);
  float v = 0.0f;
  if (s < ns) {
    if (i < n) v = sigma[(size_t)s * n + i] > 0 ? 1.0f : -1.0f;
    else if (i == n) v = 1.0f;
  }
  const uint16_t h = f2bf(v);
  A[q] = h;
  AT[(size_t)i * nsp + s] = h;
}

// T' row-major (nsp x Np) and transposed (Np x nsp)
__global__ void k_prep_T(const float* __restrict__ T, int ns, int m, int Np,
                         int nsp, float* __restrict__ Trm, float* __restrict__ Tt) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)nsp * Np) return;
  const int s = (int)(q / Np), j = (int)(q % Np);
  float v = 0.0f;
  if (s < ns) {
    if (j < m) v = T[(size_t)s * m + j];
    else if (j == m) v = 1.0f;
  }
  Trm[q] = v;
  Tt[(size_t)j * nsp + s] = v;
}

// V' (Kp x Np grid of the [W | a | b] vector p) split into three bf16 layers,
// stored transposed: vs[(layer * Np + j) * Kp + k]
__global__ void k_split_v(const float* __restrict__ p, int n, int m, int use_bias,
                          int Kp, int Np, uint16_t* __restrict__ vs) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= Np * Kp) return;
  const int j = q / Kp, k = q % Kp;
  const size_t nm = (size_t)n * m;
  float v = 0.0f;
  if (k < n && j < m) v = p[(size_t)k * m + j];
  else if (use_bias && k < n && j == m) v = p[nm + k];
  else if (use_bias && k == n && j < m) v = p[nm + n + j];
  uint16_t hi, mid, lo;
  split3(v, hi, mid, lo);
  vs[q] = hi;
  vs[(size_t)Np * Kp + q] = mid;
  vs[(size_t)2 * Np * Kp + q] = lo;
}

// y_s = sum_j T'_sj sum_k sigma'_sk V'_kj ; one wave per 16 samples
__global__ void __launch_bounds__(256)
    k_fwd_mfma(const uint16_t* __restrict__ sigA, const uint16_t* __restrict__ vs,
               const float* __restrict__ Trm, int Kp, int Np, int nsp,
               float* __restrict__ y) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int row0 = (blockIdx.x * 4 + wave) * 16;
  if (row0 >= nsp) return;
  const int nks = Kp / 32;
  const int li = l & 15, lg = l >> 4;
  bf16x8 afr[kMaxKs];
#pragma unroll
  for (int ks = 0; ks < kMaxKs; ++ks)
    if (ks < nks) afr[ks] = ld8(sigA + (size_t)(row0 + li) * Kp + ks * 32 + lg * 8);
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  const size_t layer = (size_t)Np * Kp;
  for (int ct = 0; ct < Np / 16; ++ct) {
    floatx4 c = {0.f, 0.f, 0.f, 0.f};
    const uint16_t* vb = vs + (size_t)(ct * 16 + li) * Kp + lg * 8;
#pragma unroll
    for (int sp = 0; sp < 3; ++sp)
#pragma unroll
      for (int ks = 0; ks < kMaxKs; ++ks)
        if (ks < nks)
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[ks], ld8(vb + sp * layer + ks * 32),
                                                      c, 0, 0, 0);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      acc[q] += Trm[(size_t)(row0 + 4 * lg + q) * Np + ct * 16 + li] * c[q];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int s = 8; s >= 1; s >>= 1) acc[q] += __shfl_xor(acc[q], s);
  if (li == 0)
#pragma unroll
    for (int q = 0; q < 4; ++q) y[row0 + 4 * lg + q] = acc[q];
}

// y_s <- (y_s - <omean, p>) / ns for s < ns, 0 beyond
__global__ void k_center_y(float* __restrict__ y, int ns, int nsp,
                           const double* __restrict__ dotp) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsp) return;
  const float d = (float)dotp[0];
  y[s] = s < ns ? (y[s] - d) / (float)ns : 0.0f;
}

// partial[kc][i][j] = sum_{s in chunk kc} sigma'_si y_s T'_sj ; one wave per
// (column tile, chunk); all Kp/16 row tiles accumulate in registers
__global__ void __launch_bounds__(256)
    k_bwd_mfma(const uint16_t* __restrict__ sigT, const float* __restrict__ Tt,
               const float* __restrict__ y, int Kp, int Np, int nsp,
               float* __restrict__ part) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int ct = blockIdx.x * 4 + wave;
  if (ct >= Np / 16) return;
  const int kc = blockIdx.y;
  const int li = l & 15, lg = l >> 4;
  const int nit = Kp / 16;
  floatx4 acc[kMaxIt];
#pragma unroll
  for (int it = 0; it < kMaxIt; ++it) acc[it] = (floatx4){0.f, 0.f, 0.f, 0.f};
  const float* trow = Tt + (size_t)(ct * 16 + li) * nsp;
  for (int ks = 0; ks < kBwdChunk / 32; ++ks) {
    const int s0 = kc * kBwdChunk + ks * 32 + lg * 8;
    u16x8 bh, bm, bl;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const float v = trow[s0 + e] * y[s0 + e];
      uint16_t h, md, lo;
      split3(v, h, md, lo);
      bh[e] = h;
      bm[e] = md;
      bl[e] = lo;
    }
    const bf16x8 b0 = __builtin_bit_cast(bf16x8, bh);
    const bf16x8 b1 = __builtin_bit_cast(bf16x8, bm);
    const bf16x8 b2 = __builtin_bit_cast(bf16x8, bl);
#pragma unroll
    for (int it = 0; it < kMaxIt; ++it)
      if (it < nit) {
        const bf16x8 a = ld8(sigT + (size_t)(it * 16 + li) * nsp + s0);
        acc[it] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0, acc[it], 0, 0, 0);
        acc[it] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b1, acc[it], 0, 0, 0);
        acc[it] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b2, acc[it], 0, 0, 0);
      }
  }
  float* out = part + (size_t)kc * Kp * Np;
#pragma unroll
  for (int it = 0; it < kMaxIt; ++it)
    if (it < nit)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        out[(size_t)(it * 16 + 4 * lg + q) * Np + ct * 16 + li] = acc[it][q];
}

// out[c] = G'[aug(c)] - omean_c G'[n][m] + shift p_c  (G' summed over chunks)
__global__ void k_bwd_reduce(const float* __restrict__ part, int chunks, int n, int m,
                             int use_bias, int Kp, int Np, const float* __restrict__ om,
                             double shift, const float* __restrict__ p,
                             float* __restrict__ out) {
  const size_t nm = (size_t)n * m;
  const size_t dim = use_bias ? nm + n + m : nm;
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= dim) return;
  size_t a;
  if (c < nm) a = (c / m) * Np + (c % m);
  else if (c < nm + n) a = (c - nm) * Np + m;
  else a = (size_t)n * Np + (c - nm - n);
  const size_t corner = (size_t)n * Np + m;
  const size_t stride = (size_t)Kp * Np;
  float g = 0.0f, sy = 0.0f;
  for (int k = 0; k < chunks; ++k) {
    g += part[(size_t)k * stride + a];
    sy += part[(size_t)k * stride + corner];
  }
  out[c] = (g - om[c] * sy) + (float)shift * p[c];
}

// ---- vectors and device-side CG scalars ------------------------------------
constexpr int kDotBlocks = 256;

__global__ void __launch_bounds__(256)
    k_dot32(const float* __restrict__ x, const float* __restrict__ y, size_t n,
            double* __restrict__ partial) {
  __shared__ double red[256];
  double acc = 0.0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)kDotBlocks * 256)
    acc += (double)x[i] * (double)y[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}
__global__ void __launch_bounds__(256)
    k_dot_finish(const double* __restrict__ partial, double* __restrict__ dst) {
  __shared__ double red[256];
  red[threadIdx.x] = partial[threadIdx.x];
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) dst[0] = red[0];
}
// sc layout per iteration t: sc[4t+0] = r.r (before), sc[4t+1] = p.Ap,
// sc[4t+2] = r.r (after)
__global__ void k_cg_xr(float* __restrict__ x, float* __restrict__ r,
                        const float* __restrict__ p, const float* __restrict__ Ap,
                        size_t n, const double* __restrict__ sc) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double pap = sc[1];
  const float alpha = pap != 0.0 ? (float)(sc[0] / pap) : 0.0f;
  x[i] += alpha * p[i];
  r[i] -= alpha * Ap[i];
}
__global__ void k_cg_p(float* __restrict__ p, const float* __restrict__ r, size_t n,
                       const double* __restrict__ sc) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float beta = sc[0] != 0.0 ? (float)(sc[2] / sc[0]) : 0.0f;
  p[i] = r[i] + beta * p[i];
}
__global__ void k_d2f(const double* __restrict__ a, float* __restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
__global__ void k_f2d(const float* __restrict__ a, double* __restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}
__global__ void k_copy_f(const float* __restrict__ a, float* __restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

inline unsigned nb(size_t n) { return (unsigned)((n + 255) / 256); }
inline int rup(int x, int a) { return (x + a - 1) / a * a; }

int hipfail(std::string& err, const char* what, hipError_t e) {
  err = std::string(what) + ": " + hipGetErrorString(e);
  return YQ_ERR_HIP;
}

template <typename T>
int alloc(T*& p, size_t count, std::string& err) {
  if (p) (void)hipFree(p);
  p = nullptr;
  if (hipMalloc(&p, sizeof(T) * (count ? count : 1)) != hipSuccess) {
    (void)hipGetLastError();
    err = "device allocation failed in the mixed-precision solver";
    return YQ_ERR_OOM;
  }
  return YQ_OK;
}

// dst[0] = x . y, stays on the device
void dot32(MixedWork& w, const float* x, const float* y, size_t n, double* dst,
           hipStream_t st) {
  hipLaunchKernelGGL(k_dot32, dim3(kDotBlocks), dim3(256), 0, st, x, y, n, w.dpart);
  hipLaunchKernelGGL(k_dot_finish, dim3(1), dim3(256), 0, st, w.dpart, dst);
}

}  // namespace

int mixed_prepare(MixedWork& w, const int8_t* sigma, const float* T32, int ns,
                  int n, int m, int use_bias, hipStream_t st, std::string& err) {
  if (n + 1 > 32 * kMaxKs) {
    err = "mixed solver supports n <= 159";
    return YQ_ERR_UNSUPPORTED;
  }
  w.ns = ns;
  w.n = n;
  w.m = m;
  w.use_bias = use_bias;
  w.Kp = rup(n + 1, 32);
  w.Np = rup(m + 1, 16);
  w.nsp = rup(ns, kBwdChunk);
  w.dim = (size_t)n * m + (use_bias ? (size_t)n + m : 0);
  const size_t rows = (size_t)w.nsp;
  if (rows * w.Np > w.cap_rows || w.dim > w.cap_dim) {
    const size_t chunks = rows / kBwdChunk;
    int rc;
    if ((rc = alloc(w.sigA, rows * w.Kp, err)) || (rc = alloc(w.sigT, rows * w.Kp, err)) ||
        (rc = alloc(w.Trm, rows * w.Np, err)) || (rc = alloc(w.Tt, rows * w.Np, err)) ||
        (rc = alloc(w.y, rows, err)) ||
        (rc = alloc(w.part, chunks * w.Kp * w.Np, err)) ||
        (rc = alloc(w.vsplit, (size_t)3 * w.Np * w.Kp, err)) ||
        (rc = alloc(w.om32, w.dim, err)) || (rc = alloc(w.x, w.dim, err)) ||
        (rc = alloc(w.r, w.dim, err)) || (rc = alloc(w.p, w.dim, err)) ||
        (rc = alloc(w.Ap, w.dim, err)) || (rc = alloc(w.dpart, kDotBlocks, err)))
      return rc;
    w.cap_rows = rows * w.Np;
    w.cap_dim = w.dim;
  }
  hipLaunchKernelGGL(k_prep_sig, dim3(nb(rows * w.Kp)), dim3(256), 0, st, sigma, ns, n, w.Kp,
                     w.nsp, w.sigA, w.sigT);
  hipLaunchKernelGGL(k_prep_T, dim3(nb(rows * w.Np)), dim3(256), 0, st, T32, ns, m, w.Np,
                     w.nsp, w.Trm, w.Tt);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? YQ_OK : hipfail(err, "mixed_prepare", e);
}

int mixed_apply(MixedWork& w, double shift, const float* p, float* out, hipStream_t st,
                std::string& err) {
  if (w.scal_cap < 8) {
    int rc = alloc(w.scal, 64, err);
    if (rc) return rc;
    w.scal_cap = 16;
  }
  hipLaunchKernelGGL(k_split_v, dim3(nb((size_t)w.Np * w.Kp)), dim3(256), 0, st, p, w.n, w.m,
                     w.use_bias, w.Kp, w.Np, w.vsplit);
  {
    ProfGuard g("fwd_mfma", st);
    hipLaunchKernelGGL(k_fwd_mfma, dim3((w.nsp / 16 + 3) / 4), dim3(256), 0, st, w.sigA,
                       w.vsplit, w.Trm, w.Kp, w.Np, w.nsp, w.y);
  }
  double* d = w.scal + w.scal_cap * 4 - 1;  // scratch slot for <omean, p>
  dot32(w, w.om32, p, w.dim, d, st);
  hipLaunchKernelGGL(k_center_y, dim3(nb(w.nsp)), dim3(256), 0, st, w.y, w.ns, w.nsp, d);
  const int chunks = w.nsp / kBwdChunk;
  {
    ProfGuard g("bwd_mfma", st);
    hipLaunchKernelGGL(k_bwd_mfma, dim3((w.Np / 16 + 3) / 4, chunks), dim3(256), 0, st,
                       w.sigT, w.Tt, w.y, w.Kp, w.Np, w.nsp, w.part);
  }
  ProfGuard g("bwd_reduce", st);
  hipLaunchKernelGGL(k_bwd_reduce, dim3(nb(w.dim)), dim3(256), 0, st, w.part, chunks, w.n, w.m,
                     w.use_bias, w.Kp, w.Np, w.om32, shift, p, out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? YQ_OK : hipfail(err, "mixed_apply", e);
}

void launch_mixed_apply64(MixedWork& w, double shift, const double* omean64, const double* v,
                          double* out, hipStream_t st) {
  std::string err;
  hipLaunchKernelGGL(k_d2f, dim3(nb(w.dim)), dim3(256), 0, st, omean64, w.om32, w.dim);
  hipLaunchKernelGGL(k_d2f, dim3(nb(w.dim)), dim3(256), 0, st, v, w.p, w.dim);
  (void)mixed_apply(w, shift, w.p, w.Ap, st, err);
  hipLaunchKernelGGL(k_f2d, dim3(nb(w.dim)), dim3(256), 0, st, w.Ap, out, w.dim);
}

int mixed_cg(MixedWork& w, double shift, const double* omean64, const double* b, double* x,
             double tol, int max_iter, int* iters, hipStream_t st, std::string& err) {
  const size_t dim = w.dim;
  const int need = max_iter + 2 > 16 ? max_iter + 2 : 16;
  if (w.scal_cap < need) {
    int rc = alloc(w.scal, (size_t)4 * need, err);
    if (rc) return rc;
    w.scal_cap = need;
  }
  hipLaunchKernelGGL(k_d2f, dim3(nb(dim)), dim3(256), 0, st, omean64, w.om32, dim);
  hipLaunchKernelGGL(k_d2f, dim3(nb(dim)), dim3(256), 0, st, b, w.r, dim);
  hipLaunchKernelGGL(k_copy_f, dim3(nb(dim)), dim3(256), 0, st, w.r, w.p, dim);
  hipError_t e = hipMemsetAsync(w.x, 0, sizeof(float) * dim, st);
  if (e != hipSuccess) return hipfail(err, "mixed_cg", e);
  double* sc = w.scal;
  dot32(w, w.r, w.r, dim, sc + 0, st);
  double bn2 = 0.0;
  e = hipMemcpyAsync(&bn2, sc, sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hipfail(err, "mixed_cg", e);
  *iters = 0;
  if (bn2 == 0.0) {
    hipLaunchKernelGGL(k_f2d, dim3(nb(dim)), dim3(256), 0, st, w.x, x, dim);
    return YQ_OK;
  }
  const double thr = tol * tol * bn2;
  const int check = 8;
  int it = 0;
  bool done = false;
  while (!done && it < max_iter) {
    double* s = sc + 4 * it;
    int rc = mixed_apply(w, shift, w.p, w.Ap, st, err);
    if (rc) return rc;
    dot32(w, w.p, w.Ap, dim, s + 1, st);
    hipLaunchKernelGGL(k_cg_xr, dim3(nb(dim)), dim3(256), 0, st, w.x, w.r, w.p, w.Ap, dim, s);
    dot32(w, w.r, w.r, dim, s + 2, st);
    hipLaunchKernelGGL(k_cg_p, dim3(nb(dim)), dim3(256), 0, st, w.p, w.r, dim, s);
    // next iteration's r.r before = this iteration's r.r after
    e = hipMemcpyAsync(sc + 4 * (it + 1), s + 2, sizeof(double), hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return hipfail(err, "mixed_cg", e);
    ++it;
    if (it % check == 0 || it == max_iter) {
      double rn2 = 0.0;
      e = hipMemcpyAsync(&rn2, s + 2, sizeof(double), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) return hipfail(err, "mixed_cg", e);
      if (!(rn2 >= thr)) done = true;  // converged (or NaN: stop and let FP64 check)
    }
  }
  hipLaunchKernelGGL(k_f2d, dim3(nb(dim)), dim3(256), 0, st, w.x, x, dim);
  e = hipGetLastError();
  if (e != hipSuccess) return hipfail(err, "mixed_cg", e);
  *iters = it;
  return done ? YQ_OK : YQ_ERR_NOT_CONVERGED;
}

void mixed_release(MixedWork& w) {
  for (void* p : {(void*)w.sigA, (void*)w.sigT, (void*)w.Trm, (void*)w.Tt, (void*)w.vsplit,
                  (void*)w.y, (void*)w.part, (void*)w.om32, (void*)w.x, (void*)w.r, (void*)w.p,
                  (void*)w.Ap, (void*)w.dpart, (void*)w.scal})
    if (p) (void)hipFree(p);
  w = MixedWork();
}

}  // namespace yq
