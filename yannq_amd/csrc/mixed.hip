// Mixed-precision factored SR operator on gfx950 bf16 MFMA.
//
// O's rows are sigma'_s (x) T'_s (sigma' = [sigma | 1], T' = [tanh theta | 1],
// see sr.hip).  Spins are exactly representable in bf16, so every product
// keeps float accuracy by splitting only the float operand into three bf16
// pieces (x = hi + mid + lo to 24 bits) and issuing one MFMA per piece with
// float accumulation:
//   forward   U' = sigma' V'        (nsp x Np, K = Kp)  with the fused
//             epilogue y_s = sum_j T'_sj U'_sj, split over column groups
//   backward  G' = sigma'^T diag(y) T'  (Kp x Np, K = nsp in chunks)
// v_mfma_f32_16x16x32_bf16 operand layout used here: lane l holds
// A[row l%16][k 8*(l/16) .. +7] and B[k 8*(l/16) .. +7][col l%16] and
// returns D[4*(l/16) + q][l%16], q = 0..3.
//
// The CG loop needs six launches per iteration and no host round trip: the
// vector updates carry their dot products, finished on the device by the
// last block to arrive (per-block partials summed in block order, so the
// result does not depend on scheduling).
#include <math.h>

#include "mixed.h"
#include "yq_internal.h"

namespace yq {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

constexpr int kBwdChunk = 256;  // samples per backward chunk
constexpr int kRowPad = 512;    // ns padding
constexpr int kMaxKs = 5;       // Kp <= 160 (n <= 159)
constexpr int kMaxIt = 10;      // Kp / 16
constexpr int kColTiles = 4;    // column tiles per forward group

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

// Sum v over the block (fixed tree), store the block partial, and let the
// last block to finish sum all partials in block order into *dst.
__device__ void block_dot_finish(double v, double* part, unsigned* ticket, double* dst) {
  __shared__ double red[256];
  __shared__ bool last;
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) {
    part[blockIdx.x] = red[0];
    __threadfence();
    const unsigned prev = atomicAdd(ticket, 1u);
    last = prev == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  double s = 0.0;
  for (unsigned i = t; i < gridDim.x; i += 256)
    s += __hip_atomic_load(part + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  red[t] = s;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) {
    *dst = red[0];
    *ticket = 0u;
  }
}

// index of parameter c on the (Kp x Np) grid of V' / G'
__device__ inline size_t aug_index(size_t c, int n, int m, int Np) {
  const size_t nm = (size_t)n * m;
  if (c < nm) return (c / m) * Np + (c % m);
  if (c < nm + n) return (c - nm) * Np + m;
  return (size_t)n * Np + (c - nm - n);
}

// sigma' (bf16) row-major and transposed; zero outside ns x (n+1)
__global__ void k_prep_sig(const int8_t* __restrict__ sigma, int ns, int n, int Kp, int nsp,
                           uint16_t* __restrict__ A, uint16_t* __restrict__ AT) {
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
__global__ void k_prep_T(const float* __restrict__ T, int ns, int m, int Np, int nsp,
                         float* __restrict__ Trm, float* __restrict__ Tt) {
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

// p <- r + beta p on the parameter vector (beta = rr_now / rr_prev, or 0 when
// rr_prev is null), V' = p on the (Kp x Np) grid split into three bf16 layers
// stored transposed, vs[(layer * Np + j) * Kp + k]; *dst = <omean, p>.
__global__ void __launch_bounds__(256)
    k_split_v(const float* __restrict__ r, float* __restrict__ p, const double* rr_now,
              const double* rr_prev, const float* __restrict__ om, int n, int m,
              int use_bias, int Kp, int Np, uint16_t* __restrict__ vs, double* part,
              unsigned* ticket, double* dst) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nm = (size_t)n * m;
  float v = 0.0f;
  double dv = 0.0;
  if (q < Np * Kp) {
    const int j = q / Kp, k = q % Kp;
    long c = -1;
    if (k < n && j < m) c = (long)k * m + j;
    else if (use_bias && k < n && j == m) c = (long)(nm + k);
    else if (use_bias && k == n && j < m) c = (long)(nm + n + j);
    if (c >= 0) {
      const float beta = rr_prev ? (float)(*rr_now / *rr_prev) : 0.0f;
      v = rr_prev ? r[c] + beta * p[c] : r[c];
      p[c] = v;
      dv = (double)om[c] * (double)v;
    }
    uint16_t hi, mid, lo;
    split3(v, hi, mid, lo);
    vs[q] = hi;
    vs[(size_t)Np * Kp + q] = mid;
    vs[(size_t)2 * Np * Kp + q] = lo;
  }
  block_dot_finish(dv, part, ticket, dst);
}

// ypart[cg][s] = sum over the column tiles of group cg of T'_sj U'_sj;
// one wave per 16 samples, blockIdx.y = column group
__global__ void __launch_bounds__(256)
    k_fwd_mfma(const uint16_t* __restrict__ sigA, const uint16_t* __restrict__ vs,
               const float* __restrict__ Trm, int Kp, int Np, int nsp,
               float* __restrict__ ypart) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int row0 = (blockIdx.x * 4 + wave) * 16;
  if (row0 >= nsp) return;
  const int cg = blockIdx.y;
  const int ct0 = cg * kColTiles;
  const int ct1 = min(ct0 + kColTiles, Np / 16);
  const int nks = Kp / 32;
  const int li = l & 15, lg = l >> 4;
  bf16x8 afr[kMaxKs];
#pragma unroll
  for (int ks = 0; ks < kMaxKs; ++ks)
    if (ks < nks) afr[ks] = ld8(sigA + (size_t)(row0 + li) * Kp + ks * 32 + lg * 8);
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  const size_t layer = (size_t)Np * Kp;
  for (int ct = ct0; ct < ct1; ++ct) {
    floatx4 c = {0.f, 0.f, 0.f, 0.f};
    const uint16_t* vb = vs + (size_t)(ct * 16 + li) * Kp + lg * 8;
    float tv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) tv[q] = Trm[(size_t)(row0 + 4 * lg + q) * Np + ct * 16 + li];
#pragma unroll
    for (int sp = 0; sp < 3; ++sp)
#pragma unroll
      for (int ks = 0; ks < kMaxKs; ++ks)
        if (ks < nks)
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[ks], ld8(vb + sp * layer + ks * 32),
                                                      c, 0, 0, 0);
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] += tv[q] * c[q];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int s = 8; s >= 1; s >>= 1) acc[q] += __shfl_xor(acc[q], s);
  if (li == 0)
#pragma unroll
    for (int q = 0; q < 4; ++q) ypart[(size_t)cg * nsp + row0 + 4 * lg + q] = acc[q];
}

// y_s = (sum_cg ypart - <omean, p>) / ns for s < ns, 0 beyond
__global__ void k_center_y(const float* __restrict__ ypart, int ncg, int ns, int nsp,
                           const double* __restrict__ dotp, float* __restrict__ y) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsp) return;
  float v = 0.0f;
  for (int g = 0; g < ncg; ++g) v += ypart[(size_t)g * nsp + s];
  y[s] = s < ns ? (v - (float)dotp[0]) / (float)ns : 0.0f;
}

// part[kc][i][j] = sum_{s in chunk kc} sigma'_si y_s T'_sj ; one wave per
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
#pragma unroll 2
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

// out[c] = G'[aug(c)] - omean_c G'[n][m] + shift p_c (G' summed over chunks in
// order); *dst = <p, out>
__global__ void __launch_bounds__(256)
    k_bwd_reduce(const float* __restrict__ part, int chunks, int n, int m, int use_bias,
                 int Kp, int Np, const float* __restrict__ om, double shift,
                 const float* __restrict__ p, float* __restrict__ out, double* dpart,
                 unsigned* ticket, double* dst) {
  const size_t nm = (size_t)n * m;
  const size_t dim = use_bias ? nm + n + m : nm;
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double dv = 0.0;
  if (c < dim) {
    const size_t a = aug_index(c, n, m, Np);
    const size_t corner = (size_t)n * Np + m;
    const size_t stride = (size_t)Kp * Np;
    float g = 0.0f, sy = 0.0f;
    for (int k = 0; k < chunks; ++k) {
      g += part[(size_t)k * stride + a];
      sy += part[(size_t)k * stride + corner];
    }
    const float o = (g - om[c] * sy) + (float)shift * p[c];
    out[c] = o;
    dv = (double)p[c] * (double)o;
  }
  block_dot_finish(dv, dpart, ticket, dst);
}

// x += alpha p, r -= alpha Ap with alpha = rr / pAp; *dst = <r, r>
__global__ void __launch_bounds__(256)
    k_cg_xr(float* __restrict__ x, float* __restrict__ r, const float* __restrict__ p,
            const float* __restrict__ Ap, size_t n, const double* __restrict__ sc,
            double* dpart, unsigned* ticket, double* dst) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double dv = 0.0;
  if (i < n) {
    const double pap = sc[1];
    const float alpha = pap != 0.0 ? (float)(sc[0] / pap) : 0.0f;
    x[i] += alpha * p[i];
    const float rv = r[i] - alpha * Ap[i];
    r[i] = rv;
    dv = (double)rv * (double)rv;
  }
  block_dot_finish(dv, dpart, ticket, dst);
}

__global__ void __launch_bounds__(256)
    k_dot32(const float* __restrict__ x, const float* __restrict__ y, size_t n, double* dpart,
            unsigned* ticket, double* dst) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  block_dot_finish(i < n ? (double)x[i] * (double)y[i] : 0.0, dpart, ticket, dst);
}
__global__ void k_d2f(const double* __restrict__ a, float* __restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
__global__ void k_f2d(const float* __restrict__ a, double* __restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}
__global__ void k_zero_f(float* __restrict__ a, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0f;
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

// (S + shift) p -> Ap with the products on MFMA; <omean, p> goes to dot_op and
// <p, Ap> to dot_pap.  p is first updated to r + beta p (see k_split_v).
void apply_launches(MixedWork& w, double shift, const float* r, const double* rr_now,
                    const double* rr_prev, double* dot_op, double* dot_pap, hipStream_t st) {
  const size_t grid_v = (size_t)w.Np * w.Kp;
  hipLaunchKernelGGL(k_split_v, dim3(nb(grid_v)), dim3(256), 0, st, r, w.p, rr_now, rr_prev,
                     w.om32, w.n, w.m, w.use_bias, w.Kp, w.Np, w.vsplit, w.dpart, w.ticket,
                     dot_op);
  {
    ProfGuard g("fwd_mfma", st);
    hipLaunchKernelGGL(k_fwd_mfma, dim3(w.nsp / 64, w.ncg), dim3(256), 0, st, w.sigA, w.vsplit,
                       w.Trm, w.Kp, w.Np, w.nsp, w.ypart);
  }
  hipLaunchKernelGGL(k_center_y, dim3(nb(w.nsp)), dim3(256), 0, st, w.ypart, w.ncg, w.ns, w.nsp,
                     dot_op, w.y);
  const int chunks = w.nsp / kBwdChunk;
  {
    ProfGuard g("bwd_mfma", st);
    hipLaunchKernelGGL(k_bwd_mfma, dim3((w.Np / 16 + 3) / 4, chunks), dim3(256), 0, st, w.sigT,
                       w.Tt, w.y, w.Kp, w.Np, w.nsp, w.part);
  }
  ProfGuard g("bwd_reduce", st);
  hipLaunchKernelGGL(k_bwd_reduce, dim3(nb(w.dim)), dim3(256), 0, st, w.part, chunks, w.n, w.m,
                     w.use_bias, w.Kp, w.Np, w.om32, shift, w.p, w.Ap, w.dpart, w.ticket,
                     dot_pap);
}

int ensure_scal(MixedWork& w, int iters, std::string& err) {
  const int need = iters + 4 > 16 ? iters + 4 : 16;
  if (w.scal_cap >= need) return YQ_OK;
  int rc = alloc(w.scal, (size_t)4 * need, err);
  if (rc) return rc;
  w.scal_cap = need;
  return YQ_OK;
}

}  // namespace

int mixed_prepare(MixedWork& w, const int8_t* sigma, const float* T32, int ns, int n, int m,
                  int use_bias, hipStream_t st, std::string& err) {
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
  w.nsp = rup(ns, kRowPad);
  w.ncg = (w.Np / 16 + kColTiles - 1) / kColTiles;
  w.dim = (size_t)n * m + (use_bias ? (size_t)n + m : 0);
  const size_t rows = (size_t)w.nsp;
  if (rows * w.Np > w.cap_rows || w.dim > w.cap_dim || !w.ticket) {
    const size_t chunks = rows / kBwdChunk;
    const size_t dparts = nb(w.dim > (size_t)w.Np * w.Kp ? w.dim : (size_t)w.Np * w.Kp);
    int rc;
    if ((rc = alloc(w.sigA, rows * w.Kp, err)) || (rc = alloc(w.sigT, rows * w.Kp, err)) ||
        (rc = alloc(w.Trm, rows * w.Np, err)) || (rc = alloc(w.Tt, rows * w.Np, err)) ||
        (rc = alloc(w.ypart, rows * w.ncg, err)) || (rc = alloc(w.y, rows, err)) ||
        (rc = alloc(w.part, chunks * w.Kp * w.Np, err)) ||
        (rc = alloc(w.vsplit, (size_t)3 * w.Np * w.Kp, err)) ||
        (rc = alloc(w.om32, w.dim, err)) || (rc = alloc(w.x, w.dim, err)) ||
        (rc = alloc(w.r, w.dim, err)) || (rc = alloc(w.p, w.dim, err)) ||
        (rc = alloc(w.Ap, w.dim, err)) || (rc = alloc(w.dpart, dparts, err)) ||
        (rc = alloc(w.ticket, 1, err)))
      return rc;
    hipError_t e = hipMemsetAsync(w.ticket, 0, sizeof(unsigned), st);
    if (e != hipSuccess) return hipfail(err, "mixed_prepare", e);
    w.cap_rows = rows * w.Np;
    w.cap_dim = w.dim;
  }
  hipLaunchKernelGGL(k_prep_sig, dim3(nb(rows * w.Kp)), dim3(256), 0, st, sigma, ns, n, w.Kp,
                     w.nsp, w.sigA, w.sigT);
  hipLaunchKernelGGL(k_prep_T, dim3(nb(rows * w.Np)), dim3(256), 0, st, T32, ns, m, w.Np, w.nsp,
                     w.Trm, w.Tt);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? YQ_OK : hipfail(err, "mixed_prepare", e);
}

int mixed_apply64(MixedWork& w, double shift, const double* omean64, const double* v,
                  double* out, hipStream_t st, std::string& err) {
  int rc = ensure_scal(w, 0, err);
  if (rc) return rc;
  hipLaunchKernelGGL(k_d2f, dim3(nb(w.dim)), dim3(256), 0, st, omean64, w.om32, w.dim);
  hipLaunchKernelGGL(k_d2f, dim3(nb(w.dim)), dim3(256), 0, st, v, w.r, w.dim);
  double* s = w.scal + 4 * (w.scal_cap - 1);  // scratch slot
  apply_launches(w, shift, w.r, nullptr, nullptr, s + 0, s + 1, st);
  hipLaunchKernelGGL(k_f2d, dim3(nb(w.dim)), dim3(256), 0, st, w.Ap, out, w.dim);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? YQ_OK : hipfail(err, "mixed_apply", e);
}

int mixed_cg(MixedWork& w, double shift, const double* omean64, const double* b, double* x,
             double tol, int max_iter, int* iters, hipStream_t st, std::string& err) {
  const size_t dim = w.dim;
  int rc = ensure_scal(w, max_iter, err);
  if (rc) return rc;
  hipLaunchKernelGGL(k_d2f, dim3(nb(dim)), dim3(256), 0, st, omean64, w.om32, dim);
  hipLaunchKernelGGL(k_d2f, dim3(nb(dim)), dim3(256), 0, st, b, w.r, dim);
  hipLaunchKernelGGL(k_zero_f, dim3(nb(dim)), dim3(256), 0, st, w.x, dim);
  hipLaunchKernelGGL(k_zero_f, dim3(nb(dim)), dim3(256), 0, st, w.p, dim);
  double* sc = w.scal;  // slot t: [0] r.r at the start of iteration t, [1] p.Ap, [2] <omean, p>
  hipLaunchKernelGGL(k_dot32, dim3(nb(dim)), dim3(256), 0, st, w.r, w.r, dim, w.dpart, w.ticket,
                     sc);
  double bn2 = 0.0;
  hipError_t e = hipMemcpyAsync(&bn2, sc, sizeof(double), hipMemcpyDeviceToHost, st);
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
    apply_launches(w, shift, w.r, s, it > 0 ? s - 4 : nullptr, s + 2, s + 1, st);
    hipLaunchKernelGGL(k_cg_xr, dim3(nb(dim)), dim3(256), 0, st, w.x, w.r, w.p, w.Ap, dim, s,
                       w.dpart, w.ticket, s + 4);
    ++it;
    if (it % check == 0 || it == max_iter) {
      double rn2 = 0.0;
      e = hipMemcpyAsync(&rn2, sc + 4 * it, sizeof(double), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) return hipfail(err, "mixed_cg", e);
      if (!(rn2 >= thr)) done = true;  // converged (or NaN: the FP64 check decides)
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
                  (void*)w.ypart, (void*)w.y, (void*)w.part, (void*)w.om32, (void*)w.x,
                  (void*)w.r, (void*)w.p, (void*)w.Ap, (void*)w.dpart, (void*)w.ticket,
                  (void*)w.scal})
    if (p) (void)hipFree(p);
  w = MixedWork();
}

}  // namespace yq
