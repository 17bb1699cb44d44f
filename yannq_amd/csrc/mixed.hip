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

// Two dot products finished together (partials interleaved per block).
__device__ void block_dot_finish2(double v1, double v2, double* part, unsigned* ticket,
                                  double* dst1, double* dst2) {
  __shared__ double r1[256], r2[256];
  __shared__ bool last;
  const int t = threadIdx.x;
  r1[t] = v1;
  r2[t] = v2;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (t < w) {
      r1[t] += r1[t + w];
      r2[t] += r2[t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    part[2 * blockIdx.x] = r1[0];
    part[2 * blockIdx.x + 1] = r2[0];
    __threadfence();
    const unsigned prev = atomicAdd(ticket, 1u);
    last = prev == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  double s1 = 0.0, s2 = 0.0;
  for (unsigned i = t; i < gridDim.x; i += 256) {
    s1 += __hip_atomic_load(part + 2 * i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s2 += __hip_atomic_load(part + 2 * i + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  r1[t] = s1;
  r2[t] = s2;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (t < w) {
      r1[t] += r1[t + w];
      r2[t] += r2[t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    *dst1 = r1[0];
    *dst2 = r2[0];
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

// p <- z + beta p with z = dinv * r (dinv null: z = r) and beta = rho_now /
// rho_prev (0 when rho_prev is null); V' = p on the (Kp x Np) grid split into
// three bf16 layers stored transposed, vs[(layer * Np + j) * Kp + k];
// *dst = <omean, p>.
__global__ void __launch_bounds__(256)
    k_split_v(const float* __restrict__ r, const float* __restrict__ dinv, float* __restrict__ p,
              const double* rr_now, const double* rr_prev, const float* __restrict__ om, int n,
              int m, int use_bias, int Kp, int Np, uint16_t* __restrict__ vs, double* part,
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
      const float z = dinv ? r[c] * dinv[c] : r[c];
      v = rr_prev ? z + beta * p[c] : z;
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

// ypart[cg][s] = sum over the column tiles of group cg of T'_sj U'_sj.
// Block: 4 waves x 64 rows (4 row tiles per wave); blockIdx.y = column group
// of kColTiles tiles whose three V' layers are staged once in LDS (rows padded
// by kLdsPad elements against bank conflicts) and shared by the 4 waves.
constexpr int kLdsPad = 8;
__global__ void __launch_bounds__(256)
    k_fwd_mfma(const uint16_t* __restrict__ sigA, const uint16_t* __restrict__ vs,
               const float* __restrict__ Trm, int Kp, int Np, int nsp,
               float* __restrict__ ypart) {
  extern __shared__ uint16_t bsh[];  // [3][kColTiles*16][Kp + kLdsPad]
  const int tid = threadIdx.x;
  const int cg = blockIdx.y;
  const int col0 = cg * kColTiles * 16;
  const int ncols = min(kColTiles * 16, Np - col0);
  const int ldk = Kp + kLdsPad;
  const int kchunks = Kp / 8;
  for (int q = tid; q < 3 * ncols * kchunks; q += 256) {
    const int sp = q / (ncols * kchunks);
    const int rem = q % (ncols * kchunks);
    const int col = rem / kchunks, kk = (rem % kchunks) * 8;
    *reinterpret_cast<u16x8*>(bsh + (sp * kColTiles * 16 + col) * ldk + kk) =
        *reinterpret_cast<const u16x8*>(vs + ((size_t)sp * Np + col0 + col) * Kp + kk);
  }
  __syncthreads();
  const int wave = tid >> 6, l = tid & 63;
  const int li = l & 15, lg = l >> 4;
  const int rbase = blockIdx.x * 256 + wave * 64;
  const int nks = Kp / 32;
  bf16x8 afr[4][kMaxKs];
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int ks = 0; ks < kMaxKs; ++ks)
      if (ks < nks) afr[rt][ks] = ld8(sigA + (size_t)(rbase + rt * 16 + li) * Kp + ks * 32 + lg * 8);
  float acc[4][4];
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[rt][q] = 0.f;
  for (int ct = 0; ct < ncols / 16; ++ct) {
    floatx4 c[4];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) c[rt] = (floatx4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int sp = 0; sp < 3; ++sp)
#pragma unroll
      for (int ks = 0; ks < kMaxKs; ++ks)
        if (ks < nks) {
          const bf16x8 b =
              ld8(bsh + (sp * kColTiles * 16 + ct * 16 + li) * ldk + ks * 32 + lg * 8);
#pragma unroll
          for (int rt = 0; rt < 4; ++rt)
            c[rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[rt][ks], b, c[rt], 0, 0, 0);
        }
    const int j = col0 + ct * 16 + li;
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        acc[rt][q] += Trm[(size_t)(rbase + rt * 16 + 4 * lg + q) * Np + j] * c[rt][q];
  }
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      float v = acc[rt][q];
#pragma unroll
      for (int s = 8; s >= 1; s >>= 1) v += __shfl_xor(v, s);
      if (li == 0) ypart[(size_t)cg * nsp + rbase + rt * 16 + 4 * lg + q] = v;
    }
}

// y_s = (sum_cg ypart - <omean, p>) / ns for s < ns, 0 beyond; *dst = sum_s y_s
// (the centring correction of the backward product)
__global__ void __launch_bounds__(256)
    k_center_y(const float* __restrict__ ypart, int ncg, int ns, int nsp,
               const double* __restrict__ dotp, float* __restrict__ y, double* dpart,
               unsigned* ticket, double* dst) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double dv = 0.0;
  if (s < nsp) {
    float v = 0.0f;
    for (int g = 0; g < ncg; ++g) v += ypart[(size_t)g * nsp + s];
    const float yv = s < ns ? (v - (float)dotp[0]) / (float)ns : 0.0f;
    y[s] = yv;
    dv = yv;
  }
  block_dot_finish(dv, dpart, ticket, dst);
}

// part[kc][i][j] = sum_{s in chunk kc} sigma'_si y_s T'_sj.  Block: 4 waves =
// 4 column tiles of one chunk; the chunk's sigma'^T slab is staged in LDS in
// two halves of 128 samples and shared by the waves; every wave keeps all
// Kp/16 row tiles in registers.
constexpr int kBwdHalf = 128;
__global__ void __launch_bounds__(256)
    k_bwd_mfma(const uint16_t* __restrict__ sigT, const float* __restrict__ Tt,
               const float* __restrict__ y, int Kp, int Np, int nsp,
               float* __restrict__ part) {
  __shared__ uint16_t ash[160 * (kBwdHalf + kLdsPad)];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, l = tid & 63;
  const int ct = blockIdx.x * 4 + wave;
  const bool active = ct < Np / 16;
  const int kc = blockIdx.y;
  const int li = l & 15, lg = l >> 4;
  const int nit = Kp / 16;
  const int lda = kBwdHalf + kLdsPad;
  floatx4 acc[kMaxIt];
#pragma unroll
  for (int it = 0; it < kMaxIt; ++it) acc[it] = (floatx4){0.f, 0.f, 0.f, 0.f};
  const float* trow = Tt + (size_t)(active ? ct * 16 + li : 0) * nsp;
  for (int half = 0; half < kBwdChunk / kBwdHalf; ++half) {
    const int sbase = kc * kBwdChunk + half * kBwdHalf;
    __syncthreads();
    for (int q = tid; q < Kp * (kBwdHalf / 8); q += 256) {
      const int i = q / (kBwdHalf / 8), kk = (q % (kBwdHalf / 8)) * 8;
      *reinterpret_cast<u16x8*>(ash + i * lda + kk) =
          *reinterpret_cast<const u16x8*>(sigT + (size_t)i * nsp + sbase + kk);
    }
    __syncthreads();
    if (!active) continue;
    for (int ks = 0; ks < kBwdHalf / 32; ++ks) {
      const int s0 = sbase + ks * 32 + lg * 8;
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
          const bf16x8 a = ld8(ash + (it * 16 + li) * lda + ks * 32 + lg * 8);
          acc[it] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0, acc[it], 0, 0, 0);
          acc[it] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b1, acc[it], 0, 0, 0);
          acc[it] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b2, acc[it], 0, 0, 0);
        }
    }
  }
  if (!active) return;
  float* out = part + (size_t)kc * Kp * Np;
#pragma unroll
  for (int it = 0; it < kMaxIt; ++it)
    if (it < nit)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        out[(size_t)(it * 16 + 4 * lg + q) * Np + ct * 16 + li] = acc[it][q];
}

// out[c] = G'[aug(c)] - omean_c sum(y) + shift p_c, G' summed over the chunks
// with eight interleaved partial sums combined by a fixed tree; *dst = <p, out>
__global__ void __launch_bounds__(256)
    k_bwd_reduce(const float* __restrict__ part, int chunks, int n, int m, int use_bias,
                 int Kp, int Np, const float* __restrict__ om, double shift,
                 const double* __restrict__ sum_y, const float* __restrict__ p,
                 float* __restrict__ out, double* dpart, unsigned* ticket, double* dst) {
  const size_t nm = (size_t)n * m;
  const size_t dim = use_bias ? nm + n + m : nm;
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double dv = 0.0;
  if (c < dim) {
    const size_t a = aug_index(c, n, m, Np);
    const size_t stride = (size_t)Kp * Np;
    float g8[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < chunks; k += 8) {
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (k + u < chunks) g8[u] += part[(size_t)(k + u) * stride + a];
    }
    const float g = ((g8[0] + g8[1]) + (g8[2] + g8[3])) + ((g8[4] + g8[5]) + (g8[6] + g8[7]));
    const float o = (g - om[c] * (float)sum_y[0]) + (float)shift * p[c];
    out[c] = o;
    dv = (double)p[c] * (double)o;
  }
  block_dot_finish(dv, dpart, ticket, dst);
}

// x += alpha p, r -= alpha Ap with alpha = rho / pAp; next slot: rho' = <r, dinv r>
// and <r, r>
__global__ void __launch_bounds__(256)
    k_cg_xr(float* __restrict__ x, float* __restrict__ r, const float* __restrict__ p,
            const float* __restrict__ Ap, const float* __restrict__ dinv, size_t n,
            const double* __restrict__ sc, double* dpart, unsigned* ticket, double* next) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (i < n) {
    const double pap = sc[1];
    const float alpha = pap != 0.0 ? (float)(sc[0] / pap) : 0.0f;
    x[i] += alpha * p[i];
    const float rv = r[i] - alpha * Ap[i];
    r[i] = rv;
    rr = (double)rv * (double)rv;
    rz = dinv ? (double)rv * (double)(rv * dinv[i]) : rr;
  }
  block_dot_finish2(rz, rr, dpart, ticket, next + 0, next + 4);
}

__global__ void __launch_bounds__(256)
    k_dot_rz(const float* __restrict__ r, const float* __restrict__ dinv, size_t n, double* dpart,
             unsigned* ticket, double* slot) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (i < n) {
    rr = (double)r[i] * (double)r[i];
    rz = dinv ? (double)r[i] * (double)(r[i] * dinv[i]) : rr;
  }
  block_dot_finish2(rz, rr, dpart, ticket, slot + 0, slot + 4);
}
__global__ void k_d2f(const double* __restrict__ a, float* __restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
__global__ void k_f2d(const float* __restrict__ a, double* __restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}
// Jacobi preconditioner 1 / (diag(S) + shift)
__global__ void k_jacobi_inv(const double* __restrict__ diag, double shift, float* __restrict__ out,
                             size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)(1.0 / (diag[i] + shift));
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

// (S + shift) p -> Ap with the products on MFMA; the slot s receives
// s[2] = <omean, p>, s[3] = sum(y) and s[1] = <p, Ap>.  p is first updated to
// r + beta p (see k_split_v).
void apply_launches(MixedWork& w, double shift, const float* r, const float* dinv,
                    const double* rr_now, const double* rr_prev, double* s, hipStream_t st) {
  const size_t grid_v = (size_t)w.Np * w.Kp;
  hipLaunchKernelGGL(k_split_v, dim3(nb(grid_v)), dim3(256), 0, st, r, dinv, w.p, rr_now, rr_prev,
                     w.om32, w.n, w.m, w.use_bias, w.Kp, w.Np, w.vsplit, w.dpart, w.ticket,
                     s + 2);
  {
    ProfGuard g("fwd_mfma", st);
    const size_t lds = sizeof(uint16_t) * 3 * kColTiles * 16 * (w.Kp + kLdsPad);
    hipLaunchKernelGGL(k_fwd_mfma, dim3(w.nsp / 256, w.ncg), dim3(256), lds, st, w.sigA,
                       w.vsplit, w.Trm, w.Kp, w.Np, w.nsp, w.ypart);
  }
  hipLaunchKernelGGL(k_center_y, dim3(nb(w.nsp)), dim3(256), 0, st, w.ypart, w.ncg, w.ns, w.nsp,
                     s + 2, w.y, w.dpart, w.ticket, s + 3);
  const int chunks = w.nsp / kBwdChunk;
  {
    ProfGuard g("bwd_mfma", st);
    hipLaunchKernelGGL(k_bwd_mfma, dim3((w.Np / 16 + 3) / 4, chunks), dim3(256), 0, st, w.sigT,
                       w.Tt, w.y, w.Kp, w.Np, w.nsp, w.part);
  }
  ProfGuard g("bwd_reduce", st);
  hipLaunchKernelGGL(k_bwd_reduce, dim3(nb(w.dim)), dim3(256), 0, st, w.part, chunks, w.n, w.m,
                     w.use_bias, w.Kp, w.Np, w.om32, shift, s + 3, w.p, w.Ap, w.dpart, w.ticket,
                     s + 1);
}

constexpr int kSlot = 8;  // doubles per CG iteration slot
int ensure_scal(MixedWork& w, int iters, std::string& err) {
  const int need = iters + 4 > 16 ? iters + 4 : 16;
  if (w.scal_cap >= need) return YQ_OK;
  int rc = alloc(w.scal, (size_t)kSlot * need, err);
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
        (rc = alloc(w.Ap, w.dim, err)) || (rc = alloc(w.dpart, 2 * dparts, err)) ||
        (rc = alloc(w.dinv, w.dim, err)) ||
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
  double* s = w.scal + kSlot * (w.scal_cap - 1);  // scratch slot
  apply_launches(w, shift, w.r, nullptr, nullptr, nullptr, s, st);
  hipLaunchKernelGGL(k_f2d, dim3(nb(w.dim)), dim3(256), 0, st, w.Ap, out, w.dim);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? YQ_OK : hipfail(err, "mixed_apply", e);
}

int mixed_cg(MixedWork& w, double shift, const double* omean64, const double* diag64,
             const double* b, double* x, double tol, int max_iter, int* iters, hipStream_t st,
             std::string& err) {
  const size_t dim = w.dim;
  int rc = ensure_scal(w, max_iter, err);
  if (rc) return rc;
  hipLaunchKernelGGL(k_d2f, dim3(nb(dim)), dim3(256), 0, st, omean64, w.om32, dim);
  hipLaunchKernelGGL(k_d2f, dim3(nb(dim)), dim3(256), 0, st, b, w.r, dim);
  hipLaunchKernelGGL(k_zero_f, dim3(nb(dim)), dim3(256), 0, st, w.x, dim);
  hipLaunchKernelGGL(k_zero_f, dim3(nb(dim)), dim3(256), 0, st, w.p, dim);
  const float* dinv = nullptr;
  if (diag64) {
    hipLaunchKernelGGL(k_jacobi_inv, dim3(nb(dim)), dim3(256), 0, st, diag64, shift, w.dinv, dim);
    dinv = w.dinv;
  }
  // slot t: [0] rho = <r, z>, [1] <p, Ap>, [2] <omean, p>, [3] sum y, [4] <r, r>
  double* sc = w.scal;
  hipLaunchKernelGGL(k_dot_rz, dim3(nb(dim)), dim3(256), 0, st, w.r, dinv, dim, w.dpart,
                     w.ticket, sc);
  double bn2 = 0.0;
  hipError_t e = hipMemcpyAsync(&bn2, sc + 4, sizeof(double), hipMemcpyDeviceToHost, st);
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
    double* s = sc + kSlot * it;
    apply_launches(w, shift, w.r, dinv, s, it > 0 ? s - kSlot : nullptr, s, st);
    hipLaunchKernelGGL(k_cg_xr, dim3(nb(dim)), dim3(256), 0, st, w.x, w.r, w.p, w.Ap, dinv, dim,
                       s, w.dpart, w.ticket, s + kSlot);
    ++it;
    if (it % check == 0 || it == max_iter) {
      double rn2 = 0.0;
      e = hipMemcpyAsync(&rn2, sc + kSlot * it + 4, sizeof(double), hipMemcpyDeviceToHost, st);
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
                  (void*)w.r, (void*)w.p, (void*)w.Ap, (void*)w.dpart, (void*)w.ticket, (void*)w.dinv,
                  (void*)w.scal})
    if (p) (void)hipFree(p);
  w = MixedWork();
}

}  // namespace yq
