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
// The CG loop (Chronopoulos-Gear form) needs three launches per iteration
// (the fused reduction + vector update + split, and the two products) and no
// host round trip: the kernels carry their dot products as per-block
// partials, which the next kernel that needs the value sums itself in a fixed
// order.
#include <math.h>

#include <algorithm>
#include <array>
#include <cstring>
#include <map>
#include <mutex>

#include "mixed.h"
#include "yq_internal.h"

namespace yq {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

constexpr int kBwdChunkMin = 128;  // smallest backward chunk (sizes the partials)
constexpr int kRowPad = 512;    // ns padding
constexpr int kMaxKs = 5;       // register-resident k-steps (Kp <= 160); past that the
                                // chunked forward and the row-grouped backward run
constexpr int kColTilesMin = 2;  // narrowest forward column group (sizes ypart)

// device control block of the CG loop: ints [stop flag, completed iterations,
// iteration budget, outcome (kConvYes: <r, r> below the threshold when the
// loop stopped, kConvNonFinite: <b, b> or <r, r> was NaN or infinite, 0:
// neither), chunks published to the host]; the first four are what the host
// polls
enum { kCtlStop = 0, kCtlIters = 1, kCtlMaxIt = 2, kCtlConv = 3, kCtlChunk = 4, kCtlInts = 8 };
enum { kConvYes = 1, kConvNonFinite = 2 };
constexpr int kSlotInts = 8;  // one host slot: 4 control ints, <r, r>, threshold
// and doubles [shift, stopping threshold on <r, r>, alpha x 2, gamma x 2, the
// latest <r, r> (what the host's convergence forecast reads)]
enum { kCtlShift = 0, kCtlThr = 1, kCtlAlpha = 2, kCtlGamma = 4, kCtlRr = 6, kCtlDoubles = 7 };

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
// x = hi + mid + lo exactly: each piece takes the next 8 significant bits
// (truncation: the top 16 bits of a float are a bf16, and each remainder is
// exact -- x has 24 significant bits, hi, mid and lo 8 each).  Five VALU
// operations a value, against some sixteen for round-to-nearest pieces,
// which would be no more exact.
__device__ inline void split3(float x, uint16_t& hi, uint16_t& mid, uint16_t& lo) {
  const uint32_t u = __float_as_uint(x);
  const float r1 = x - __uint_as_float(u & 0xFFFF0000u);
  const uint32_t u1 = __float_as_uint(r1);
  const float r2 = r1 - __uint_as_float(u1 & 0xFFFF0000u);
  hi = (uint16_t)(u >> 16);
  mid = (uint16_t)(u1 >> 16);
  lo = (uint16_t)(__float_as_uint(r2) >> 16);
}
__device__ inline bf16x8 ld8(const uint16_t* p) {
  const u16x8 v = *reinterpret_cast<const u16x8*>(p);
  return __builtin_bit_cast(bf16x8, v);
}

// Dot products travel between launches as per-block partial sums: the
// producing kernel stores one double per block (fixed tree over the block),
// and every wave of a consuming kernel sums the partials itself in a fixed
// order (lane-strided sums, then an xor butterfly).  No fences, atomics or
// counters are involved, and the value does not depend on scheduling.
// Fixed-order block sum of NV doubles over the first nwaves waves of the
// block (xor butterfly over the lanes, then the waves in index order);
// thread 0 stores total k to out[k][slot].  Every thread of the block must
// call it (one barrier inside).
template <int NV>
__device__ void block_sums(const double (&v)[NV], double* const (&out)[NV], int slot,
                           int nwaves) {
  __shared__ double red[NV][16];
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double x = v[k];
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) x += __shfl_xor(x, w);
    if (lane == 0 && wave < nwaves) red[k][wave] = x;
  }
  __syncthreads();
  if (t == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double x = 0.0;
      for (int w = 0; w < nwaves; ++w) x += red[k][w];
      out[k][slot] = x;
    }
}
// total of cnt partials, identical in every lane of every wave: lane l adds
// partials l, l + 64, ... in order (loaded 8 at a time so that the loads are
// in flight together), then an xor butterfly over the wave
__device__ inline double psum(const double* __restrict__ part, int cnt) {
  const int l = threadIdx.x & 63;
  double s = 0.0;
  for (int i0 = l; i0 < cnt; i0 += 512) {
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = i0 + 64 * q < cnt ? part[i0 + 64 * q] : 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (i0 + 64 * q < cnt) s += v[q];
  }
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) s += __shfl_xor(s, w);
  return s;
}
// psum in two parts, so that other loads can be in flight with the partials:
// psum_issue loads this lane's first 8 (partials l, l + 64, ..., l + 448),
// psum_finish adds them and the rest in psum's order
// (cnt >= 1; the loads are unconditional at clamped indices, the unused ones
// are left out of the sum)
__device__ inline void psum_issue(const double* __restrict__ part, int cnt, double (&v)[8]) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < 8; ++q) v[q] = part[min(l + 64 * q, cnt - 1)];
}
__device__ inline double psum_finish(const double* __restrict__ part, int cnt,
                                     const double (&v)[8]) {
  const int l = threadIdx.x & 63;
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (l + 64 * q < cnt) s += v[q];
  for (int i0 = l + 512; i0 < cnt; i0 += 512) {
    double w[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) w[q] = i0 + 64 * q < cnt ? part[i0 + 64 * q] : 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (i0 + 64 * q < cnt) s += w[q];
  }
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) s += __shfl_xor(s, w);
  return s;
}
// three totals at once, each in the order of psum
__device__ inline void psum3(const double* __restrict__ a, int na, const double* __restrict__ b,
                             int nb, const double* __restrict__ c, int nc, double& sa, double& sb,
                             double& sc) {
  const int l = threadIdx.x & 63;
  const int top = max(na, max(nb, nc));
  double x = 0.0, y = 0.0, z = 0.0;
  for (int i0 = l; i0 < top; i0 += 256) {
    double va[4], vb[4], vc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = i0 + 64 * q;
      va[q] = i < na ? a[i] : 0.0;
      vb[q] = i < nb ? b[i] : 0.0;
      vc[q] = i < nc ? c[i] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = i0 + 64 * q;
      if (i < na) x += va[q];
      if (i < nb) y += vb[q];
      if (i < nc) z += vc[q];
    }
  }
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) {
    x += __shfl_xor(x, w);
    y += __shfl_xor(y, w);
    z += __shfl_xor(z, w);
  }
  sa = x;
  sb = y;
  sc = z;
}

// index of parameter c on the (Kp x Np) grid of V' / G'
__device__ inline size_t aug_index(size_t c, int n, int m, int Np) {
  const size_t nm = (size_t)n * m;
  if (c < nm) return (c / m) * Np + (c % m);
  if (c < nm + n) return (c - nm) * Np + m;
  return (size_t)n * Np + (c - nm - n);
}

// sigma' as bits, a set bit standing for -1 and a clear one for +1: row-major
// (nsp rows of kSigRow bytes; bit k of row s is bit k % 8 of byte k / 8) and
// transposed (Kp rows of nsp / 8 bytes; bit s of row k is bit s % 8 of byte
// s / 8).  Column n (the bias) is +1, and so are the padding columns k > n
// and the padding rows s >= ns: the rows k > n of V' are zero and those of G'
// are never read, and a padded sample has T' = 0 and y = 0, so their signs do
// not matter.  A kernel reads 1 bit per spin instead of a bf16 and expands
// the bits into bf16 +-1 in registers (expand8).
constexpr int kSigRow = 32;  // bytes per row-major bit row up to Kp = 256
// bytes of a row-major bit row for Kp (a multiple of 16, at least kSigRow)
inline int sig_row_bytes(int Kp) { return std::max(kSigRow, (Kp / 8 + 15) / 16 * 16); }
__global__ void k_prep_sig_rows(const int8_t* __restrict__ sigma, int ns, int n, int nsp,
                                int row_bytes, uint8_t* __restrict__ A) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)nsp * row_bytes) return;
  const int s = (int)(q / row_bytes), b = (int)(q % row_bytes);
  unsigned v = 0;
  if (s < ns)
    for (int e = 0; e < 8; ++e) {
      const int k = 8 * b + e;
      if (k < n && sigma[(size_t)s * n + k] < 0) v |= 1u << e;
    }
  A[q] = (uint8_t)v;
}
__global__ void k_prep_sig_cols(const int8_t* __restrict__ sigma, int ns, int n, int Kp, int nsp,
                                uint8_t* __restrict__ T) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int rb = nsp / 8;
  if (q >= (size_t)Kp * rb) return;
  const int k = (int)(q / rb), sb = (int)(q % rb);
  unsigned v = 0;
  if (k < n)
    for (int e = 0; e < 8; ++e) {
      const int s = 8 * sb + e;
      if (s < ns && sigma[(size_t)s * n + k] < 0) v |= 1u << e;
    }
  T[q] = (uint8_t)v;
}
// 8 spins (bits 0..7 of b, set = -1) as 8 bf16 (+1 = 0x3F80, -1 = 0xBF80):
// bf16 pair p of the result takes bits 2p (low half) and 2p + 1 (high half)
// into the sign positions 15 and 31
__device__ inline u16x8 expand8(uint32_t b) {
  typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
  u32x4 w;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const uint32_t x = b >> (2 * p);
    w[p] = 0x3F803F80u | ((x & 1u) << 15) | ((x & 2u) << 30);
  }
  return __builtin_bit_cast(u16x8, w);
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

// ypart[cg][s] = sum over the column tiles of group cg of T'_sj U'_sj.
// The product is formed transposed, U'^T = V'^T sigma'^T (A = V'^T from LDS,
// B = sigma'^T: the 8 consecutive k of one sample), so a lane's four results
// belong to one sample and four consecutive columns: the T' epilogue is one
// 16-byte load and one accumulator per sample tile, folded over the four lane
// groups at the end.  Block: 4 waves x 64 samples (4 sample tiles per wave);
// blockIdx.y = column group of CT tiles whose three V' layers are
// staged once in LDS (rows padded by kLdsPad elements against bank
// conflicts) and shared by the 4 waves.  NKS = Kp / 32 k-steps (the sigma'
// fragments live in registers).  The block is latency-bound, so every
// global load is issued as early as the registers allow: the T' of up to 4
// column tiles is requested before the products that need it.
constexpr int kLdsPad = 8;
// (bx, by): the block's place in the machine's own (nsp / 256) x ncg grid
template <int CT, int NKS>
__device__ __forceinline__ void fwd_body(int bx, int by, const uint8_t* __restrict__ sigA,
                                         const uint16_t* __restrict__ vs,
                                         const float* __restrict__ Trm, int Kp, int Np, int nsp,
                                         float* __restrict__ ypart, int* __restrict__ ictl,
                                         int count) {
  // fused multiply-adds in the products (their sums need no replay; the
  // split of bwd's y T' is exact either way, every piece a truncation)
  #pragma clang fp contract(fast)
  extern __shared__ uint16_t bsh[];  // [3][CT*16][Kp + kLdsPad]
  const int tid = threadIdx.x;
  const int cg = by;
  const int col0 = cg * CT * 16;
  const int ncols = min(CT * 16, Np - col0);
  const int ldk = Kp + kLdsPad;
  const int kchunks = Kp / 8;
  const int wave = tid >> 6, l = tid & 63;
  const int li = l & 15, lg = l >> 4;
  const int rbase = bx * 256 + wave * 64;
  // The V' slab goes through registers into LDS (at most 16 16-byte pieces
  // per thread, the LDS budget of the slab).  All of the block's loads are
  // independent of one another: the slab, this wave's sigma' fragments and
  // the T' of the first column tiles are issued together before the first
  // wait (every block of a launch is resident at once, so the registers this
  // takes cost no concurrency).
  constexpr int kStage = (3 * CT * 16 * 4 * NKS + 255) / 256 < 16
                             ? (3 * CT * 16 * 4 * NKS + 255) / 256 : 16;
  const int total = 3 * ncols * kchunks;
  // piece q = tid + 256 i is (layer sp, column col, k chunk kc8) with
  // q = (sp * ncols + col) * kchunks + kc8; the decomposition advances by 256
  // per piece with one division at the start
  const int d_col = 256 / kchunks, d_kc = 256 % kchunks;
  int sp0 = tid / (ncols * kchunks);
  int col0_ = (tid / kchunks) % ncols, kc0 = tid % kchunks;
  auto advance = [&](int& sp, int& col, int& kc8) {
    kc8 += d_kc;
    col += d_col;
    if (kc8 >= kchunks) {
      kc8 -= kchunks;
      ++col;
    }
    while (col >= ncols) {
      col -= ncols;
      ++sp;
    }
  };
  u16x8 stg[kStage];
  {
    int sp = sp0, col = col0_, kc8 = kc0;
#pragma unroll
    for (int i = 0; i < kStage; ++i) {
      if (tid + 256 * i < total)
        stg[i] = *reinterpret_cast<const u16x8*>(vs + ((size_t)sp * Np + col0 + col) * Kp +
                                                 kc8 * 8);
      advance(sp, col, kc8);
    }
  }
  // this lane's spin row of each sample tile: the k-step ks fragment is
  // byte 4 ks + lg of the row (bits of k = 32 ks + 8 lg .. + 7), i.e. byte
  // lg of dword ks
  typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
  constexpr int kSegs = (NKS + 3) / 4;  // 16-byte pieces of the row in use
  u32x4 sbits[4][kSegs];
#pragma unroll
  for (int st = 0; st < 4; ++st)
#pragma unroll
    for (int h = 0; h < kSegs; ++h)
      sbits[st][h] = *reinterpret_cast<const u32x4*>(
          sigA + (size_t)(rbase + st * 16 + li) * kSigRow + 16 * h);
  constexpr int PF = CT < 4 ? CT : 4;  // column tiles whose T' is in flight together
  const int nct = ncols / 16;
  floatx4 t[PF][4];
  auto load_t = [&](int cb) {
#pragma unroll
    for (int q = 0; q < PF; ++q)
      if (cb + q < nct) {
        const int j0 = col0 + (cb + q) * 16 + 4 * lg;
#pragma unroll
        for (int st = 0; st < 4; ++st)
          t[q][st] =
              *reinterpret_cast<const floatx4*>(Trm + (size_t)(rbase + st * 16 + li) * Np + j0);
      }
  };
  load_t(0);
  // The stop flag is tested once the data loads are in flight (the buffers
  // of a stopped run stay valid, so the loads are harmless); the empty asm
  // keeps the compiler from sinking the loads below the test.
  asm volatile("" ::: "memory");
  if (ictl[kCtlStop]) return;
  // count: this product belongs to a CG iteration whose update ran (one
  // writer; the next update reads the counter)
  if (count && bx == 0 && by == 0 && tid == 0) ictl[kCtlIters] += 1;
  {
    int sp = sp0, col = col0_, kc8 = kc0;
#pragma unroll
    for (int i = 0; i < kStage; ++i) {
      if (tid + 256 * i < total)
        *reinterpret_cast<u16x8*>(bsh + (sp * CT * 16 + col) * ldk + kc8 * 8) = stg[i];
      advance(sp, col, kc8);
    }
  }
  __syncthreads();
  bf16x8 sfr[4][NKS];
#pragma unroll
  for (int st = 0; st < 4; ++st)
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks)
      sfr[st][ks] = __builtin_bit_cast(
          bf16x8, expand8((sbits[st][ks >> 2][ks & 3] >> (8 * lg)) & 0xFFu));
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  for (int cb = 0; cb < nct; cb += PF) {
    if (cb > 0) load_t(cb);
#pragma unroll
    for (int q = 0; q < PF; ++q) {
      const int ct = cb + q;
      if (ct < nct) {
        floatx4 c[4];
#pragma unroll
        for (int st = 0; st < 4; ++st) c[st] = (floatx4){0.f, 0.f, 0.f, 0.f};
        // the column tile's V' operands are read from LDS together, then used
        // (read next to their products, every read would wait on its own)
        bf16x8 vv[3][NKS];
#pragma unroll
        for (int sp = 0; sp < 3; ++sp)
#pragma unroll
          for (int ks = 0; ks < NKS; ++ks)
            vv[sp][ks] = ld8(bsh + (sp * CT * 16 + ct * 16 + li) * ldk + ks * 32 + lg * 8);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int sp = 0; sp < 3; ++sp)
#pragma unroll
          for (int ks = 0; ks < NKS; ++ks)
#pragma unroll
            for (int st = 0; st < 4; ++st)
              c[st] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vv[sp][ks], sfr[st][ks], c[st], 0,
                                                              0, 0);
        // c[st][e] = U'^T[col0 + ct*16 + 4*lg + e][rbase + st*16 + li]
#pragma unroll
        for (int st = 0; st < 4; ++st)
          acc[st] += ((t[q][st][0] * c[st][0] + t[q][st][1] * c[st][1]) +
                      (t[q][st][2] * c[st][2] + t[q][st][3] * c[st][3]));
      }
    }
  }
#pragma unroll
  for (int st = 0; st < 4; ++st) {
    float v = acc[st];
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    if (lg == 0) ypart[(size_t)cg * nsp + rbase + st * 16 + li] = v;
  }
}
template <int CT, int NKS>
__global__ void __launch_bounds__(256)
    k_fwd_mfma(const uint8_t* __restrict__ sigA, const uint16_t* __restrict__ vs,
               const float* __restrict__ Trm, int Kp, int Np, int nsp,
               float* __restrict__ ypart, int* __restrict__ ictl, int count) {
  fwd_body<CT, NKS>(blockIdx.x, blockIdx.y, sigA, vs, Trm, Kp, Np, nsp, ypart, ictl, count);
}

// forward kernel for (CT, Kp / 32)
template <int CT>
auto fwd_kernel_ks(int nks) {
  switch (nks) {
    case 1: return k_fwd_mfma<CT, 1>;
    case 2: return k_fwd_mfma<CT, 2>;
    case 3: return k_fwd_mfma<CT, 3>;
    case 4: return k_fwd_mfma<CT, 4>;
    default: return k_fwd_mfma<CT, 5>;
  }
}
auto fwd_kernel(int ct, int nks) {
  return ct == 2 ? fwd_kernel_ks<2>(nks) : ct == 8 ? fwd_kernel_ks<8>(nks) : fwd_kernel_ks<4>(nks);
}

// Forward product past Kp = 32 kMaxKs (n >= 160): K in chunks of kFwdKc
// k-steps, each chunk's V' rows staged through LDS for the column group, the
// U' accumulators of the CT column tiles kept in registers across the
// chunks, and this lane's spin fragments read per chunk (16 bytes of the
// sig_row-byte bit row).  Same block shape, operand layouts and epilogue as
// fwd_body; the sums over k run chunk by chunk (layer, then k-step, within a
// chunk).
constexpr int kFwdKc = 4;
template <int CT>
__global__ void __launch_bounds__(256)
    k_fwd_mfma_g(const uint8_t* __restrict__ sigA, int sig_row, const uint16_t* __restrict__ vs,
                 const float* __restrict__ Trm, int Kp, int Np, int nsp,
                 float* __restrict__ ypart, int* __restrict__ ictl, int count) {
  // fused multiply-adds in the products (their sums need no replay; the
  // split of bwd's y T' is exact either way, every piece a truncation)
  #pragma clang fp contract(fast)
  constexpr int ldk = 32 * kFwdKc + kLdsPad;
  extern __shared__ uint16_t bsh[];  // [3][CT*16][ldk]
  const int tid = threadIdx.x;
  const int wave = tid >> 6, l = tid & 63;
  const int li = l & 15, lg = l >> 4;
  const int cg = blockIdx.y;
  const int col0 = cg * CT * 16;
  const int ncols = min(CT * 16, Np - col0);
  const int nct = ncols / 16;
  const int rbase = blockIdx.x * 256 + wave * 64;
  if (ictl[kCtlStop]) return;  // uniform per launch, before any barrier
  if (count && blockIdx.x == 0 && blockIdx.y == 0 && tid == 0) ictl[kCtlIters] += 1;
  const int nks = Kp / 32;
  floatx4 c[CT][4];
#pragma unroll
  for (int ct = 0; ct < CT; ++ct)
#pragma unroll
    for (int st = 0; st < 4; ++st) c[ct][st] = (floatx4){0.f, 0.f, 0.f, 0.f};
  typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
  for (int k0 = 0; k0 < nks; k0 += kFwdKc) {
    const int kn = min(kFwdKc, nks - k0);  // k-steps in this chunk
    __syncthreads();  // the previous chunk's operands are read
    // V' rows [32 k0, 32 (k0 + kn)) of the group's columns, in 8-k pieces
    const int kch = 4 * kn;
    const int total = 3 * ncols * kch;
    for (int q = tid; q < total; q += 256) {
      const int kc8 = q % kch, col = (q / kch) % ncols, sp = q / (kch * ncols);
      *reinterpret_cast<u16x8*>(bsh + (sp * CT * 16 + col) * ldk + kc8 * 8) =
          *reinterpret_cast<const u16x8*>(vs + ((size_t)sp * Np + col0 + col) * Kp + k0 * 32 +
                                          kc8 * 8);
    }
    // k-step k0 + ks of this lane's row: byte lg of dword ks of the row's
    // 16-byte piece at byte 4 k0
    bf16x8 sfr[4][kFwdKc];
#pragma unroll
    for (int st = 0; st < 4; ++st) {
      const u32x4 w4 = *reinterpret_cast<const u32x4*>(
          sigA + (size_t)(rbase + st * 16 + li) * sig_row + 4 * k0);
#pragma unroll
      for (int ks = 0; ks < kFwdKc; ++ks)
        sfr[st][ks] = __builtin_bit_cast(bf16x8, expand8((w4[ks] >> (8 * lg)) & 0xFFu));
    }
    __syncthreads();
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
      if (ct >= nct) break;
      bf16x8 vv[3][kFwdKc];
#pragma unroll
      for (int sp = 0; sp < 3; ++sp)
#pragma unroll
        for (int ks = 0; ks < kFwdKc; ++ks)
          vv[sp][ks] = ld8(bsh + (sp * CT * 16 + ct * 16 + li) * ldk + ks * 32 + lg * 8);
#pragma unroll
      for (int sp = 0; sp < 3; ++sp)
#pragma unroll
        for (int ks = 0; ks < kFwdKc; ++ks) {
          if (ks >= kn) break;  // uniform
#pragma unroll
          for (int st = 0; st < 4; ++st)
            c[ct][st] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vv[sp][ks], sfr[st][ks],
                                                                c[ct][st], 0, 0, 0);
        }
    }
  }
  // c[ct][st][e] = U'^T[col0 + ct*16 + 4*lg + e][rbase + st*16 + li]
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ct = 0; ct < CT; ++ct) {
    if (ct >= nct) break;
    const int j0 = col0 + ct * 16 + 4 * lg;
#pragma unroll
    for (int st = 0; st < 4; ++st) {
      const floatx4 t =
          *reinterpret_cast<const floatx4*>(Trm + (size_t)(rbase + st * 16 + li) * Np + j0);
      acc[st] += ((t[0] * c[ct][st][0] + t[1] * c[ct][st][1]) +
                  (t[2] * c[ct][st][2] + t[3] * c[ct][st][3]));
    }
  }
#pragma unroll
  for (int st = 0; st < 4; ++st) {
    float v = acc[st];
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    if (lg == 0) ypart[(size_t)cg * nsp + rbase + st * 16 + li] = v;
  }
}
// k width of the forward's LDS slab: all of Kp, or one chunk past kMaxKs
inline int fwd_k_width(int Kp) { return Kp <= 32 * kMaxKs ? Kp : 32 * kFwdKc; }

// part[kc][i][j] = sum_{s in chunk kc} sigma'_si y_s T'_sj with the centred
// forward result y_s = (sum_cg ypart - <omean, p>) / ns (0 for padded rows),
// formed in LDS for the chunk; blocks of column group 0 also store the chunk
// sums of y (ys, the centring correction of the reduction) and of y^2 (yy:
// <u, S u> = ns sum_s y_s^2, the CG curvature).  Block: 4 waves =
// 4 column tiles of one chunk; the chunk's sigma'^T slab is staged in LDS in
// slices of SLICE samples (sized by Kp, so the LDS per block stays small
// and several blocks share a CU) and shared by the waves; every wave keeps
// all Kp/16 row tiles in registers.

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2): hardware block b runs on XCD b % 8.  Renumbering so that XCD x takes
// the consecutive logical blocks [x * per, (x + 1) * per) keeps the column
// groups of one chunk on one XCD, where they share its sigma'^T slab in L2.
// A bijection on [0, total); the tail past a multiple of 8 keeps its index.
constexpr int kXcds = 8;
__device__ inline int xcd_block(int b, int total) {
  const int per = total / kXcds;
  if (b >= per * kXcds) return b;
  return (b % kXcds) * per + b / kXcds;
}

// (b, gx, total): the block's linear index in the machine's own
// ((Np / 16 + 3) / 4) x chunks grid, the grid's x extent and its size
// NIT = Kp / 16 row tiles (a template parameter, so that the row-tile loop
// unrolls without branches and the LDS reads of its A operands are issued
// together).  RG (Kp > 32 kMaxKs): the grid is ((Np / 16 + 3) / 4) x chunks x
// row groups, each block forming rows [rg NIT 16, (rg + 1) NIT 16) of G'
// (the last group may be shorter), with only those rows' sigma'^T bits in LDS.
template <int CHUNK, int SLICE, int NIT, bool RG = false>
__device__ __forceinline__ void bwd_body(int b, int gx, int total,
                                         const uint8_t* __restrict__ sigT,
                                         const float* __restrict__ Tt,
                                         const float* __restrict__ ypart, int ncg, int ns,
                                         const double* __restrict__ op, int nop, int Kp, int Np,
                                         int nsp, float* __restrict__ part,
                                         double* __restrict__ ys, double* __restrict__ yy,
                                         const int* __restrict__ ictl) {
  // fused multiply-adds in the products (their sums need no replay; the
  // split of bwd's y T' is exact either way, every piece a truncation)
  #pragma clang fp contract(fast)
  extern __shared__ uint16_t ash[];  // [Kp][SLICE + kLdsPad]
  __shared__ float ysh[CHUNK];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, l = tid & 63;
  const int lb = xcd_block(b, total);
  const int cx = lb % gx;
  const int kc = RG ? (lb / gx) % (nsp / CHUNK) : lb / gx;
  const int rg = RG ? (lb / gx) / (nsp / CHUNK) : 0;
  const int r0 = rg * NIT * 16;                         // first row of G' formed here
  const int nrow = RG ? min(NIT * 16, Kp - r0) : Kp;   // rows formed here
  const int ct = cx * 4 + wave;
  const bool active = ct < Np / 16;
  const int li = l & 15, lg = l >> 4;
  const int lda = SLICE + kLdsPad;
  const float* trow = Tt + (size_t)(active ? ct * 16 + li : 0) * nsp;
  // Software pipeline over the slices: the sigma'^T bits (32 samples of one
  // row per word) and this lane's T' values of the next kDepth slices are in
  // flight while the products of a slice run.  The bits are expanded to bf16
  // on their way into LDS (Kp <= 160 bounds the words per thread).
  constexpr int kWpr = SLICE / 32;  // words of a row per slice
  constexpr int kSlab = (32 * kMaxKs * kWpr + 255) / 256;
  constexpr int kKs = SLICE / 32;  // k-steps of 32 samples per slice
  constexpr int kHalves = CHUNK / SLICE;
  constexpr int kDepth = kHalves < 2 ? kHalves : 2;  // slices in flight
  const int nslab = nrow * kWpr;
  const int rbytes = nsp / 8;  // bytes of a transposed bit row
  uint32_t sg[kDepth][kSlab];
  floatx4 tv[kDepth][2 * kKs];  // T'[ct*16 + li][sbase + ks*32 + lg*8 + 0..7]
  // (slot: a compile-time index once the slice loop is unrolled)
  auto load_slice = [&](int half, int slot) {
    const int sbase = kc * CHUNK + half * SLICE;
#pragma unroll
    for (int i = 0; i < kSlab; ++i) {
      const int q = tid + 256 * i;
      if (q < nslab) {
        const int row = q / kWpr, wd = q % kWpr;
        sg[slot][i] = *reinterpret_cast<const uint32_t*>(sigT + (size_t)(r0 + row) * rbytes +
                                                         sbase / 8 + 4 * wd);
      }
    }
#pragma unroll
    for (int ks = 0; ks < kKs; ++ks) {  // (an inactive wave reads row 0: in bounds, unused)
      const float* tp = trow + sbase + ks * 32 + lg * 8;
      tv[slot][2 * ks] = *reinterpret_cast<const floatx4*>(tp);
      tv[slot][2 * ks + 1] = *reinterpret_cast<const floatx4*>(tp + 4);
    }
  };
  // Every independent load is issued before the first wait: the first
  // kDepth slices, the first 8 column groups' forward partials of this
  // thread's samples and the <omean, u> partials; then the stop flag.
#pragma unroll
  for (int h = 0; h < kDepth; ++h) load_slice(h, h);
  constexpr int kYr = (CHUNK + 255) / 256;  // samples of the chunk per thread
  // (clamped indices instead of conditions: a conditional load whose value
  // is selected at once makes the wave wait for it)
  float yl[kYr][8];
#pragma unroll
  for (int r = 0; r < kYr; ++r) {
    const int q = min(tid + 256 * r, CHUNK - 1);
#pragma unroll
    for (int e = 0; e < 8; ++e)
      yl[r][e] = ypart[(size_t)min(e, ncg - 1) * nsp + kc * CHUNK + q];
  }
  double opv[8];
  psum_issue(op, nop, opv);
  // the stop flag is tested after the y values (a test before them would let
  // the compiler move these loads behind it)
  const int stop = ictl[kCtlStop];
  {
    const float ov = (float)psum_finish(op, nop, opv);
    double dv = 0.0, dv2 = 0.0;
#pragma unroll
    for (int r = 0; r < kYr; ++r) {
      const int q = tid + 256 * r;
      if (q >= CHUNK) break;
      const int s = kc * CHUNK + q;
      float v = 0.0f;
#pragma unroll
      for (int e = 0; e < 8; ++e)
        if (e < ncg) v += yl[r][e];
      for (int g0 = 8; g0 < ncg; g0 += 8) {  // loads in flight together, sums in order
        float y8[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) y8[e] = g0 + e < ncg ? ypart[(size_t)(g0 + e) * nsp + s] : 0.0f;
#pragma unroll
        for (int e = 0; e < 8; ++e)
          if (g0 + e < ncg) v += y8[e];
      }
      const float yv = s < ns ? (v - ov) / (float)ns : 0.0f;
      ysh[q] = yv;
      dv += (double)yv;
      dv2 += (double)yv * (double)yv;
    }
    if (stop) return;  // uniform per launch, before any barrier or global store
    if (cx == 0 && rg == 0) {  // uniform per block
      const double v2[2] = {dv, dv2};
      double* const o2[2] = {ys, yy};
      block_sums<2>(v2, o2, kc, 4);
    }
  }
  floatx4 acc[NIT];
#pragma unroll
  for (int it = 0; it < NIT; ++it) acc[it] = (floatx4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int half = 0; half < kHalves; ++half) {
    const int sbase = kc * CHUNK + half * SLICE;
    const int slot = half % kDepth;
    __syncthreads();  // the previous slice's products are done with the LDS
#pragma unroll
    for (int i = 0; i < kSlab; ++i) {
      const int q = tid + 256 * i;
      if (q < nslab) {
        const int row = q / kWpr, wd = q % kWpr;
#pragma unroll
        for (int j = 0; j < 4; ++j)
          *reinterpret_cast<u16x8*>(ash + row * lda + 32 * wd + 8 * j) =
              expand8((sg[slot][i] >> (8 * j)) & 0xFFu);
      }
    }
    floatx4 tc[2 * kKs];
#pragma unroll
    for (int i = 0; i < 2 * kKs; ++i) tc[i] = tv[slot][i];
    if (half + kDepth < kHalves) load_slice(half + kDepth, slot);
    __syncthreads();
    if (!active) continue;
#pragma unroll
    for (int ks = 0; ks < kKs; ++ks) {
      const int s0 = sbase + ks * 32 + lg * 8;
      // the k-step's A operands are read from LDS first, so that their
      // latency passes under the split below (read one at a time next to
      // their products, each would wait on its own)
      bf16x8 av[NIT];
#pragma unroll
      for (int it = 0; it < NIT; ++it)
        if (!RG || it * 16 < nrow) av[it] = ld8(ash + (it * 16 + li) * lda + ks * 32 + lg * 8);
      __builtin_amdgcn_sched_barrier(0);  // (the scheduler would sink them again)
      u16x8 bh, bm, bl;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const float v = tc[2 * ks + (e >> 2)][e & 3] * ysh[s0 - kc * CHUNK + e];
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
      for (int it = 0; it < NIT; ++it) {
        if (RG && it * 16 >= nrow) break;  // uniform
        acc[it] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[it], b0, acc[it], 0, 0, 0);
        acc[it] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[it], b1, acc[it], 0, 0, 0);
        acc[it] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[it], b2, acc[it], 0, 0, 0);
      }
    }
  }
  if (!active) return;
  float* out = part + (size_t)kc * Kp * Np;
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    if (RG && it * 16 >= nrow) break;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      out[(size_t)(r0 + it * 16 + 4 * lg + q) * Np + ct * 16 + li] = acc[it][q];
  }
}
template <int CHUNK, int SLICE, int NIT, bool RG = false>
__global__ void __launch_bounds__(256)
    k_bwd_mfma(const uint8_t* __restrict__ sigT, const float* __restrict__ Tt,
               const float* __restrict__ ypart, int ncg, int ns, const double* __restrict__ op,
               int nop, int Kp, int Np, int nsp, float* __restrict__ part,
               double* __restrict__ ys, double* __restrict__ yy, const int* __restrict__ ictl) {
  // grid (x, chunks x row groups): the linear index runs over all of it
  bwd_body<CHUNK, SLICE, NIT, RG>(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x,
                                  gridDim.x * gridDim.y, sigT, Tt, ypart, ncg, ns, op, nop, Kp,
                                  Np, nsp, part, ys, yy, ictl);
}

// backward kernel for a chunk length; slices of 64 samples (longer slices,
// with fewer latency steps but less overlap, measured slower at n = 100,
// m = 400, 16384 samples)
constexpr int kBwdSlice = 64;
template <int CHUNK>
auto bwd_kernel_it(int nit) {
  switch (nit) {
    case 1: return k_bwd_mfma<CHUNK, kBwdSlice, 1>;
    case 2: return k_bwd_mfma<CHUNK, kBwdSlice, 2>;
    case 3: return k_bwd_mfma<CHUNK, kBwdSlice, 3>;
    case 4: return k_bwd_mfma<CHUNK, kBwdSlice, 4>;
    case 5: return k_bwd_mfma<CHUNK, kBwdSlice, 5>;
    case 6: return k_bwd_mfma<CHUNK, kBwdSlice, 6>;
    case 7: return k_bwd_mfma<CHUNK, kBwdSlice, 7>;
    case 8: return k_bwd_mfma<CHUNK, kBwdSlice, 8>;
    case 9: return k_bwd_mfma<CHUNK, kBwdSlice, 9>;
    default: return k_bwd_mfma<CHUNK, kBwdSlice, 10>;
  }
}
// (Kr = n + 1 rounded up to 16, at most 160 here: nit = Kr / 16 is 1 .. 10)
auto bwd_kernel(int chunk, int nit) {
  return chunk == 128 ? bwd_kernel_it<128>(nit)
                      : chunk == 512 ? bwd_kernel_it<512>(nit) : bwd_kernel_it<256>(nit);
}
// past Kp = 160: row groups of kBwdGroupTiles tiles
constexpr int kBwdGroupTiles = 2 * kMaxKs;
auto bwd_kernel_rg(int chunk) {
  return chunk == 128   ? k_bwd_mfma<128, kBwdSlice, kBwdGroupTiles, true>
         : chunk == 512 ? k_bwd_mfma<512, kBwdSlice, kBwdGroupTiles, true>
                        : k_bwd_mfma<256, kBwdSlice, kBwdGroupTiles, true>;
}

// One Chronopoulos-Gear CG step for (S + shift) x = b with preconditioner M
// (M = diag(dinv), or the identity when dinv is null), fused with the
// reduction of the previous product and the bf16 split of the vector the
// next product needs.  With u = M r, w = (S + shift) u, gamma = <r, u> and
// delta = <w, u> = ns sum_s y_s^2 + shift <u, u> (partials of earlier
// launches):
//   mode 0 (start):   u = M r
//   mode 1 (first):   alpha = gamma / delta, p = u, s = w
//   mode 2:           beta = gamma / gamma_prev,
//                     alpha = gamma / (delta - beta gamma / alpha_prev),
//                     p = u + beta p, s = w + beta s
//   mode 3 (product): out = w only
// then (modes 1, 2) x += alpha p, r -= alpha s, u = M r.  w is formed here
// from the backward chunk partials, w_c = G'[aug(c)] - omean_c sum(y) +
// shift u_c.  The new u is stored and split into V' on the (Kp x Np) grid in
// three bf16 layers stored column-major, vs[(layer * Np + j) * Kp + k]; the
// partials of <omean, u>, <r, u>, <r, r> and <u, u> of the new vectors go to
// op, g_out, rr_out and uu_out.  The predicate and the step sizes read the
// slots of parity par and write those of par ^ 1; alpha and gamma of this
// iteration are kept in control-block slot par for the next one.
// Block: 1024 threads on a 16 x 16 tile of the grid.  Four groups of 256
// threads sum the chunk partials g, g + 4, ... of the tile (8 loads in flight
// per thread; combined as (g0 + g1) + (g2 + g3)); group 0 then updates the
// vectors, read along j (consecutive parameters), and writes V' along k
// through an LDS transpose.
// (b, gx, total): the block's linear index in the machine's own
// (Np / 16) x (Kp / 16) grid, the grid's x extent and its size
__device__ __forceinline__ void cg_step_body(int b, int gx, int total, float* __restrict__ x, float* __restrict__ r, float* __restrict__ p,
              float* __restrict__ s, float* __restrict__ u, float* __restrict__ out,
              const float* __restrict__ dinv, int mode, int par,
              const double* __restrict__ g_in, const double* __restrict__ rr_in,
              const double* __restrict__ uu_in, int nv, const double* __restrict__ ys,
              const double* __restrict__ yy, int nys, int ns, double* __restrict__ g_out,
              double* __restrict__ rr_out, double* __restrict__ uu_out, double* __restrict__ op,
              double* dctl, int* ictl, const float* __restrict__ part, int chunks,
              const float* __restrict__ om, int n, int m, int use_bias, int Kp, int Np,
              uint16_t* __restrict__ vs) {
  __shared__ float grp[4][256];
  __shared__ uint16_t tile[3][16][17];
  const int t = threadIdx.x, g = t >> 8, e = t & 255;
  const int G = blockDim.x >> 8;  // thread groups of 256 (1, 2 or 4)
  // a tile row of the chunk partials is 64 bytes, half an L2 line: the
  // XCD-aware numbering puts tiles (2i, 2i + 1) of a row on one XCD so that
  // the line is fetched from HBM once
  const int lb = xcd_block(b, total);
  const int j0 = (lb % gx) * 16, k0 = (lb / gx) * 16;
  const size_t nm = (size_t)n * m;
  const int kk = e >> 4, jj = e & 15;
  const int k = k0 + kk, j = j0 + jj;
  long c = -1;
  if (k < n && j < m) c = (long)k * m + j;
  else if (use_bias && k < n && j == m) c = (long)(nm + k);
  else if (use_bias && k == n && j < m) c = (long)(nm + n + j);
  // the control block, read before anything waits (only block 0 writes it,
  // into slots that no block reads in this launch)
  const double shift = dctl[kCtlShift];
  const double thr = dctl[kCtlThr];
  const double gam_prev = dctl[kCtlGamma + (par ^ 1)];
  const double alpha_prev = dctl[kCtlAlpha + (par ^ 1)];
  const int stop0 = ictl[kCtlStop];
  const int iters = ictl[kCtlIters];
  const int max_it = ictl[kCtlMaxIt];
  // Every independent load is issued before the first wait: this thread's
  // chunk partials (up to 16, G == 4) and, in group 0, its vector elements;
  // the values are used in the order of the code below.
  constexpr int kPre = 16;
  float pre[kPre];
  // the chunk partials hold the grid's rows (Kr = n + 1 rounded up to 16)
  const size_t pstride = (size_t)((total / gx) * 16) * Np;
  const size_t paug = (size_t)k * Np + j;  // aug(c)
  if (mode != 0 && c >= 0 && G == 4) {  // clamped chunk index: unconditional loads
#pragma unroll
    for (int q = 0; q < kPre; ++q) pre[q] = part[(size_t)min(g + 4 * q, chunks - 1) * pstride + paug];
  }
  float r_c = 0.0f, u_c = 0.0f, p_c = 0.0f, s_c = 0.0f, x_c = 0.0f, om_c = 0.0f, di_c = 1.0f;
  if (g == 0 && c >= 0) {
    r_c = r[c];
    u_c = u[c];
    om_c = om[c];
    if (mode == 2) {
      p_c = p[c];
      s_c = s[c];
    }
    if (mode == 1 || mode == 2) x_c = x[c];
    if (dinv) di_c = dinv[c];
  }
  asm volatile("" ::: "memory");  // the loads stay above the first wait
  float alpha = 0.0f, beta = 0.0f;
  // The grid totals of the partials (<r, r>, <r, u>, <u, u>, sum y^2, sum y)
  // are summed by wave 0 alone and shared through LDS: every wave would get
  // the same values, and nv partials read by each of the 16 waves of every
  // block would cost 16 x 3 x nv^2 doubles of L2 reads per launch.
  __shared__ double tot[5];
  if (mode != 0) {
    if (t < 64) {
      double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      if (mode == 1 || mode == 2) {
        psum3(rr_in, nv, g_in, nv, uu_in, nv, v[0], v[1], v[2]);
        v[3] = psum(yy, nys);
      }
      v[4] = psum(ys, nys);
      if (t == 0)
#pragma unroll
        for (int q = 0; q < 5; ++q) tot[q] = v[q];
    }
    __syncthreads();
  }
  if (mode == 1 || mode == 2) {
    const double rn2 = tot[0], gam = tot[1], uu = tot[2];
    if (lb == 0 && t == 0) dctl[kCtlRr] = rn2;
    // CG continuation predicate (uniform over the grid): stop once <r, r>
    // falls below the threshold, is not finite, or the iteration budget is
    // spent; the later kernels of the iteration see the flag and return at
    // once
    const bool fin = __builtin_isfinite(rn2);
    const bool conv = fin && rn2 < thr;
    if (stop0 || conv || !fin || iters >= max_it) {
      if (lb == 0 && t == 0 && !stop0) {
        ictl[kCtlStop] = 1;
        ictl[kCtlConv] = conv ? kConvYes : fin ? 0 : kConvNonFinite;
      }
      return;
    }
    const double del = (double)ns * tot[3] + shift * uu;
    double a, bt = 0.0;
    if (mode == 1) {
      a = gam / del;
    } else {
      bt = gam / gam_prev;
      a = gam / (del - bt * gam / alpha_prev);
    }
    if (!__builtin_isfinite(a)) a = 0.0;  // breakdown (delta or its correction 0): no step
    alpha = (float)a;
    beta = (float)bt;
    if (lb == 0 && t == 0) {
      dctl[kCtlAlpha + par] = a;
      dctl[kCtlGamma + par] = gam;
    }
  }
  float wv = 0.0f;
  if (mode != 0) {
    // The chunk partials of this element fall into the residue classes
    // r = 0..3 of the chunk index, each summed in chunk order and combined as
    // (a0 + a1) + (a2 + a3); thread group g sums the classes r = g, g + G, ...
    // (up to 8 chunks of each class in flight), so the value does not depend
    // on G.
    if (c >= 0 && G == 4) {  // one class per group
      const size_t a = paug;
      const size_t stride = pstride;
      // (starts from the first partial rather than 0 + it, which would make
      // the compiler wait for that load where it is issued)
      float acc = pre[0];
#pragma unroll
      for (int q = 1; q < kPre; ++q)
        if (g + 4 * q < chunks) acc += pre[q];
      for (int q0 = g + 4 * kPre; q0 < chunks; q0 += 32) {
        float v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
          v[q] = q0 + 4 * q < chunks ? part[(size_t)(q0 + 4 * q) * stride + a] : 0.0f;
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (q0 + 4 * q < chunks) acc += v[q];
      }
      grp[g][e] = acc;
    } else if (c >= 0) {
      const size_t a = paug;
      const size_t stride = pstride;
      float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      for (int q0 = 0; q0 < chunks; q0 += 32) {
        float v[4][8];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int ch = q0 + r + 4 * q;
            v[r][q] = (r & (G - 1)) == g && ch < chunks ? part[(size_t)ch * stride + a] : 0.0f;
          }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int q = 0; q < 8; ++q)
            if ((r & (G - 1)) == g && q0 + r + 4 * q < chunks) acc[r] += v[r][q];
      }
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if ((r & (G - 1)) == g) grp[r][e] = acc[r];
    }
    const float sy = (float)tot[4];
    __syncthreads();
    if (g == 0 && c >= 0) {
      const float gs = (grp[0][e] + grp[1][e]) + (grp[2][e] + grp[3][e]);
      wv = (gs - om_c * sy) + (float)shift * u_c;
    }
    if (mode == 3) {
      if (g == 0 && c >= 0) out[c] = wv;
      return;  // uniform per launch
    }
  }
  double d_om = 0.0, d_g = 0.0, d_rr = 0.0, d_uu = 0.0;
  if (g == 0) {
    float v = 0.0f;
    if (c >= 0) {
      float rv = r_c;
      if (mode > 0) {
        const float uo = u_c;
        const float pv = mode == 1 ? uo : uo + beta * p_c;
        const float sv = mode == 1 ? wv : wv + beta * s_c;
        p[c] = pv;
        s[c] = sv;
        x[c] = x_c + alpha * pv;
        rv = rv - alpha * sv;
        r[c] = rv;
      }
      v = dinv ? rv * di_c : rv;
      u[c] = v;
      d_om = (double)om_c * (double)v;
      d_rr = (double)rv * (double)rv;
      d_g = (double)rv * (double)v;
      d_uu = (double)v * (double)v;
    }
    uint16_t hi, mid, lo;
    split3(v, hi, mid, lo);
    tile[0][jj][kk] = hi;
    tile[1][jj][kk] = mid;
    tile[2][jj][kk] = lo;
  }
  {
    const double v4[4] = {d_om, d_g, d_rr, d_uu};
    double* const o4[4] = {op, g_out, rr_out, uu_out};
    block_sums<4>(v4, o4, lb, 4);  // barrier inside
  }
  if (g == 0) {
    const int tj = e >> 4, tk = e & 15;
#pragma unroll
    for (int sp = 0; sp < 3; ++sp)
      vs[((size_t)sp * Np + j0 + tj) * Kp + k0 + tk] = tile[sp][tj][tk];
  }
}
__global__ void __launch_bounds__(1024)
    k_cg_step(float* __restrict__ x, float* __restrict__ r, float* __restrict__ p,
              float* __restrict__ s, float* __restrict__ u, float* __restrict__ out,
              const float* __restrict__ dinv, int mode, int par,
              const double* __restrict__ g_in, const double* __restrict__ rr_in,
              const double* __restrict__ uu_in, int nv, const double* __restrict__ ys,
              const double* __restrict__ yy, int nys, int ns, double* __restrict__ g_out,
              double* __restrict__ rr_out, double* __restrict__ uu_out, double* __restrict__ op,
              double* dctl, int* ictl, const float* __restrict__ part, int chunks,
              const float* __restrict__ om, int n, int m, int use_bias, int Kp, int Np,
              uint16_t* __restrict__ vs) {
  cg_step_body(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x, gridDim.x * gridDim.y,
               x, r, p, s, u, out, dinv, mode, par, g_in, rr_in, uu_in, nv, ys, yy, nys, ns, g_out, rr_out, uu_out, op, dctl, ictl, part, chunks, om, n, m, use_bias, Kp, Np, vs);
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
// control block of a CG run (a kernel, so the reset is also valid in a graph)
__global__ void k_set_ctl(double* dctl, int* ictl, double shift, double thr, int max_it) {
  if (threadIdx.x == 0) {
    dctl[kCtlShift] = shift;
    dctl[kCtlThr] = thr;
    ictl[kCtlStop] = 0;
    ictl[kCtlIters] = 0;
    ictl[kCtlMaxIt] = max_it;
    ictl[kCtlConv] = 0;
    ictl[kCtlChunk] = 0;
  }
}
// Start of a CG run without a host round trip: <b, b> = the sum of the <r, r>
// partials of slot 0 in block order (staged in LDS, added by one thread: the
// order of a host loop), then the threshold tol^2 <b, b> and the budget;
// b = 0 stops the run at once (x stays 0, reported as converged), a NaN or
// infinite <b, b> too (reported as non-finite).
__global__ void __launch_bounds__(256)
    k_start_rr(const double* __restrict__ rr0, int nbv, double tol, int max_it, double* dctl,
               int* ictl) {
  __shared__ double sh[1024];
  double bn2 = 0.0;
  for (int base = 0; base < nbv; base += 1024) {
    const int cnt = min(1024, nbv - base);
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) sh[i] = rr0[base + i];
    __syncthreads();
    if (threadIdx.x == 0)
      for (int i = 0; i < cnt; ++i) bn2 += sh[i];
  }
  if (threadIdx.x == 0) {
    ictl[kCtlIters] = 0;
    if (bn2 == 0.0 || !__builtin_isfinite(bn2)) {
      ictl[kCtlStop] = 1;
      ictl[kCtlConv] = bn2 == 0.0 ? kConvYes : kConvNonFinite;
    } else {
      dctl[kCtlThr] = tol * tol * bn2;
      ictl[kCtlStop] = 0;
      ictl[kCtlConv] = 0;
      ictl[kCtlMaxIt] = max_it;
    }
  }
}
// The control block's first four ints, the latest <r, r> and the threshold
// into host slot (chunk count) & 1 of the pinned mirror (8 ints a slot: the
// ints, then the two doubles): the last node of every chunk of iterations, so
// the host reads the state after chunk k in slot k & 1 once the chunk's event
// has completed (the event's completion makes the stores visible to the host).
// A run that spent its budget before any k_cg_step tested the residual of its
// last update (the budget ends on a chunk boundary) is tested here, on the
// <r, r> partials that update left (slot iters & 1, rr0 / rr1), summed in the
// order k_cg_step sums them.  One wave.
__global__ void __launch_bounds__(64)
    k_publish(int* ictl, double* __restrict__ dctl, int* host_ctl,
              const double* __restrict__ rr0, const double* __restrict__ rr1, int nbv) {
  if (!ictl[kCtlStop] && ictl[kCtlIters] >= ictl[kCtlMaxIt]) {  // uniform
    const double rn2 = psum((ictl[kCtlIters] & 1) ? rr1 : rr0, nbv);
    if (threadIdx.x == 0) {
      const bool fin = __builtin_isfinite(rn2);
      dctl[kCtlRr] = rn2;
      ictl[kCtlStop] = 1;
      ictl[kCtlConv] = fin && rn2 < dctl[kCtlThr] ? kConvYes : fin ? 0 : kConvNonFinite;
    }
  }
  if (threadIdx.x == 0) {
    const int k = ictl[kCtlChunk];
    ictl[kCtlChunk] = k + 1;
    int* o = host_ctl + kSlotInts * (k & 1);
#pragma unroll
    for (int i = 0; i < 4; ++i) o[i] = ictl[i];
    double* od = reinterpret_cast<double*>(o + 4);
    od[0] = dctl[kCtlRr];
    od[1] = dctl[kCtlThr];
  }
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

// block counts of the partial-sum producers and their offsets in w.red:
// gamma = <r, u>, <r, r> and <u, u> alternate between two slots (iteration
// parity)
struct RedLayout {
  int nbv, nby;
  size_t g[2], rr[2], uu[2], op, ys, yy, total;
};
RedLayout red_layout(const MixedWork& w, int chunk = 0) {
  RedLayout L;
  if (chunk == 0) chunk = w.chunk;
  L.nbv = (w.Kr / 16) * (w.Np / 16);         // k_cg_step tiles
  L.nby = w.nsp / chunk;                      // k_bwd_mfma chunks
  size_t o = 0;
  for (int q = 0; q < 2; ++q) {
    L.g[q] = o;
    L.rr[q] = o + L.nbv;
    L.uu[q] = o + 2 * L.nbv;
    o += 3 * L.nbv;
  }
  L.op = o;
  L.ys = L.op + L.nbv;
  L.yy = L.ys + L.nby;
  L.total = L.yy + L.nby;
  return L;
}

void publish_launch(MixedWork& w, hipStream_t st) {
  const RedLayout L = red_layout(w);
  hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, st, w.ictl, w.dctl, w.host_ctl_dev,
                     w.red + L.rr[0], w.red + L.rr[1], L.nbv);
}

// the two products for the u last split into w.vsplit by k_cg_step, leaving
// the backward chunk partials (reduced by the next k_cg_step); every kernel
// returns at once when the stop flag is set.  count: the product belongs to a
// CG iteration.
void product_launches(MixedWork& w, int count, hipStream_t st) {
  const RedLayout L = red_layout(w);
  const size_t lds_f = sizeof(uint16_t) * 3 * w.ct * 16 * (fwd_k_width(w.Kp) + kLdsPad);
  const int chunks = w.nsp / w.chunk;
  if (w.Kp <= 32 * kMaxKs) {
    launch_prof("fwd_mfma", fwd_kernel(w.ct, w.Kp / 32), dim3(w.nsp / 256, w.ncg), dim3(256),
                lds_f, st, w.sigA, w.vsplit, w.Trm, w.Kp, w.Np, w.nsp, w.ypart, w.ictl, count);
    const size_t lds = sizeof(uint16_t) * w.Kr * (kBwdSlice + kLdsPad);
    launch_prof("bwd_mfma", bwd_kernel(w.chunk, w.Kr / 16), dim3((w.Np / 16 + 3) / 4, chunks),
                dim3(256), lds, st, w.sigT, w.Tt, w.ypart, w.ncg, w.ns, w.red + L.op, L.nbv,
                w.Kr, w.Np, w.nsp, w.part, w.red + L.ys, w.red + L.yy, w.ictl);
    return;
  }
  const int srow = sig_row_bytes(w.Kp);
  launch_prof("fwd_mfma", w.ct == 2 ? k_fwd_mfma_g<2> : k_fwd_mfma_g<4>,
              dim3(w.nsp / 256, w.ncg), dim3(256), lds_f, st, w.sigA, srow, w.vsplit, w.Trm,
              w.Kp, w.Np, w.nsp, w.ypart, w.ictl, count);
  const int groups = (w.Kr / 16 + kBwdGroupTiles - 1) / kBwdGroupTiles;
  const size_t lds = sizeof(uint16_t) * 16 * kBwdGroupTiles * (kBwdSlice + kLdsPad);
  launch_prof("bwd_mfma", bwd_kernel_rg(w.chunk),
              dim3((w.Np / 16 + 3) / 4, chunks * groups), dim3(256), lds, st, w.sigT, w.Tt,
              w.ypart, w.ncg, w.ns, w.red + L.op, L.nbv, w.Kr, w.Np, w.nsp, w.part,
              w.red + L.ys, w.red + L.yy, w.ictl);
}

// thread groups of 256 per k_cg_step block for a number of chunk partials
// (any count gives the same values; fewer, smaller blocks when there are few)
int cg_groups(int chunks) { return chunks > 16 ? 4 : chunks > 8 ? 2 : 1; }

// k_cg_step in the given mode for iteration parity par (mode 0 writes slot 0)
void step_launch(MixedWork& w, const float* dinv, int mode, int par, hipStream_t st) {
  const RedLayout L = red_layout(w);
  const int wr = mode == 0 ? 0 : par ^ 1;
  launch_prof("cg_step", k_cg_step, dim3(w.Np / 16, w.Kr / 16),
              dim3(256 * cg_groups(w.nsp / w.chunk)), 0, st, w.x, w.r,
              w.p, w.s, w.u, w.Ap, dinv, mode, par, w.red + L.g[par], w.red + L.rr[par],
              w.red + L.uu[par], L.nbv, w.red + L.ys, w.red + L.yy, L.nby, w.ns,
              w.red + L.g[wr], w.red + L.rr[wr], w.red + L.uu[wr], w.red + L.op, w.dctl, w.ictl,
              w.part, w.nsp / w.chunk, w.om32, w.n, w.m, w.use_bias, w.Kp, w.Np, w.vsplit);
}

// w.Ap = (S + shift) v for v in w.r, without preconditioning: the split
// (mode 0), the products and the reduction (mode 3)
void apply_launches(MixedWork& w, hipStream_t st) {
  step_launch(w, nullptr, 0, 0, st);
  product_launches(w, 0, st);
  step_launch(w, nullptr, 3, 0, st);
}

// one CG iteration (parity par): the step, then the products on the new u
void iteration_launches(MixedWork& w, const float* dinv, int par, bool first, hipStream_t st) {
  step_launch(w, dinv, first ? 1 : 2, par, st);
  product_launches(w, 1, st);
}

// ---------------------------------------------------------------- grouped CG
// The inner CG of several machines in one set of launches per iteration: each
// grouped kernel's grid is the concatenation of the machines' own grids, a
// block table maps a block to (machine, block of the machine's own grid), and
// the block runs that machine's kernel body with the machine's own tile
// shapes, so every machine's arithmetic is that of its own launches.  A
// machine that is not iterating has its stop flag set and its blocks return
// at once.
struct GroupDesc {
  const uint8_t* sigA;
  const uint8_t* sigT;
  const float* Trm;
  const float* Tt;
  const float* dinv;     // w.dinv when preconditioned, else null
  const double* om64;    // FP64 column mean of O
  const double* b64;     // right-hand side
  const double* diag64;  // diag S (Jacobi), or null
  double* x64;           // solution
  uint16_t* vs;
  float *ypart, *part, *om, *x, *r, *p, *s, *u, *Ap, *dinv_w;
  double *red, *dctl;
  int* ictl;
  long long dim;
  long long g[2], rr[2], uu[2], op, ys, yy;  // offsets into red (red_layout)
  double tol;            // relative tolerance of this call
  int ns, n, m, use_bias, Kp, Kr, Np, nsp, ncg, ct, chunk, nbv, nby;
  int fwd_gx, bwd_gx, bwd_total, cg_gx, cg_total;
  int in;                // takes part in this call
  int max_it;            // iteration budget of this call
};

template <int CT>
__device__ __forceinline__ void fwd_dispatch(int nks, int bx, int by, const GroupDesc& d,
                                             int count) {
  switch (nks) {
    case 1: fwd_body<CT, 1>(bx, by, d.sigA, d.vs, d.Trm, d.Kp, d.Np, d.nsp, d.ypart, d.ictl, count); break;
    case 2: fwd_body<CT, 2>(bx, by, d.sigA, d.vs, d.Trm, d.Kp, d.Np, d.nsp, d.ypart, d.ictl, count); break;
    case 3: fwd_body<CT, 3>(bx, by, d.sigA, d.vs, d.Trm, d.Kp, d.Np, d.nsp, d.ypart, d.ictl, count); break;
    case 4: fwd_body<CT, 4>(bx, by, d.sigA, d.vs, d.Trm, d.Kp, d.Np, d.nsp, d.ypart, d.ictl, count); break;
    default: fwd_body<CT, 5>(bx, by, d.sigA, d.vs, d.Trm, d.Kp, d.Np, d.nsp, d.ypart, d.ictl, count); break;
  }
}
// bt[block] = (machine, linear block of its (nsp / 256) x ncg grid)
__global__ void __launch_bounds__(256)
    k_fwd_group(const GroupDesc* __restrict__ gd, const int2* __restrict__ bt, int count) {
  const int2 e = bt[blockIdx.x];
  const GroupDesc& d = gd[e.x];
  const int bx = e.y % d.fwd_gx, by = e.y / d.fwd_gx;
  const int nks = d.Kp / 32;
  if (d.ct == 2) fwd_dispatch<2>(nks, bx, by, d, count);
  else if (d.ct == 8) fwd_dispatch<8>(nks, bx, by, d, count);
  else fwd_dispatch<4>(nks, bx, by, d, count);
}
// bt[block] = (machine, linear block of its ((Np / 16 + 3) / 4) x chunks grid)
__global__ void __launch_bounds__(256)
    k_bwd_group(const GroupDesc* __restrict__ gd, const int2* __restrict__ bt) {
  const int2 e = bt[blockIdx.x];
  const GroupDesc& d = gd[e.x];
#define YQ_BWD_ARGS                                                                            \
  e.y, d.bwd_gx, d.bwd_total, d.sigT, d.Tt, d.ypart, d.ncg, d.ns, d.red + d.op, d.nbv, d.Kr,   \
      d.Np, d.nsp, d.part, d.red + d.ys, d.red + d.yy, d.ictl
#define YQ_BWD_NIT(CH)                                                                         \
  switch (d.Kr / 16) {                                                                         \
    case 1: bwd_body<CH, kBwdSlice, 1>(YQ_BWD_ARGS); break;                                    \
    case 2: bwd_body<CH, kBwdSlice, 2>(YQ_BWD_ARGS); break;                                    \
    case 3: bwd_body<CH, kBwdSlice, 3>(YQ_BWD_ARGS); break;                                    \
    case 4: bwd_body<CH, kBwdSlice, 4>(YQ_BWD_ARGS); break;                                    \
    case 5: bwd_body<CH, kBwdSlice, 5>(YQ_BWD_ARGS); break;                                    \
    case 6: bwd_body<CH, kBwdSlice, 6>(YQ_BWD_ARGS); break;                                    \
    case 7: bwd_body<CH, kBwdSlice, 7>(YQ_BWD_ARGS); break;                                    \
    case 8: bwd_body<CH, kBwdSlice, 8>(YQ_BWD_ARGS); break;                                    \
    case 9: bwd_body<CH, kBwdSlice, 9>(YQ_BWD_ARGS); break;                                    \
    default: bwd_body<CH, kBwdSlice, 10>(YQ_BWD_ARGS); break;                                  \
  }
  if (d.chunk == 128) {
    YQ_BWD_NIT(128)
  } else if (d.chunk == 512) {
    YQ_BWD_NIT(512)
  } else {
    YQ_BWD_NIT(256)
  }
#undef YQ_BWD_NIT
#undef YQ_BWD_ARGS
}
// bt[block] = (machine, linear block of its (Np / 16) x (Kp / 16) grid)
__global__ void __launch_bounds__(1024)
    k_cg_group(const GroupDesc* __restrict__ gd, const int2* __restrict__ bt, int mode, int par) {
  const int2 e = bt[blockIdx.x];
  const GroupDesc& d = gd[e.x];
  if (d.ictl[kCtlStop]) return;  // not iterating (uniform per block)
  const int wr = mode == 0 ? 0 : par ^ 1;
  double* red = d.red;
  cg_step_body(e.y, d.cg_gx, d.cg_total, d.x, d.r, d.p, d.s, d.u, d.Ap, d.dinv, mode, par,
               red + d.g[par], red + d.rr[par], red + d.uu[par], d.nbv, red + d.ys, red + d.yy,
               d.nby, d.ns, red + d.g[wr], red + d.rr[wr], red + d.uu[wr], red + d.op, d.dctl,
               d.ictl, d.part, d.nsp / d.chunk, d.om, d.n, d.m, d.use_bias, d.Kp, d.Np, d.vs);
}
// start of a grouped call (blockIdx.y = machine): the float copies of the
// column mean and of b, x = 0, the Jacobi inverse, and the control block
// reset as for a product (machines not taking part: stop flag set)
__global__ void k_group_start(const GroupDesc* __restrict__ gd, double shift) {
  const GroupDesc& d = gd[blockIdx.y];
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (!d.in) {
    if (lead) d.ictl[kCtlStop] = 1;
    return;
  }
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < d.dim) {
    d.om[i] = (float)d.om64[i];
    d.r[i] = (float)d.b64[i];
    d.x[i] = 0.0f;
    if (d.diag64) d.dinv_w[i] = (float)(1.0 / (d.diag64[i] + shift));
  }
  if (lead) {
    d.dctl[kCtlShift] = shift;
    d.dctl[kCtlThr] = 0.0;
    d.ictl[kCtlStop] = 0;
    d.ictl[kCtlIters] = 0;
    d.ictl[kCtlMaxIt] = 1;
  }
}
// sum of the <r, r> partials of slot sl in block order, as the host sums them
__device__ inline double rr_total(const GroupDesc& d, int sl) {
  const double* a = d.red + d.rr[sl];
  double s = 0.0;
  for (int i0 = 0; i0 < d.nbv; i0 += 16) {
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = i0 + q < d.nbv ? a[i0 + q] : 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (i0 + q < d.nbv) s += v[q];
  }
  return s;
}
// one thread per machine.  which 0, after the start products: stat[q] =
// <b, b>, and the control block of the iterations (threshold tol^2 <b, b>,
// budget; b = 0 stops at once with x = 0).  which 1, at the end: stat[count +
// q] = <r, r> after the completed iterations, stat[2 count + q] = their count.
__global__ void k_group_rr(const GroupDesc* __restrict__ gd, int count, int which,
                           double* __restrict__ stat) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= count) return;
  const GroupDesc& d = gd[q];
  if (!d.in) return;
  if (which == 0) {
    const double bn2 = rr_total(d, 0);
    stat[q] = bn2;
    if (bn2 == 0.0 || !__builtin_isfinite(bn2)) {
      d.ictl[kCtlStop] = 1;
    } else {
      d.dctl[kCtlThr] = d.tol * d.tol * bn2;
      d.ictl[kCtlStop] = 0;
      d.ictl[kCtlIters] = 0;
      d.ictl[kCtlMaxIt] = d.max_it;
    }
  } else {
    const int it = d.ictl[kCtlIters];
    stat[count + q] = rr_total(d, it & 1);
    stat[2 * count + q] = (double)it;
  }
}
// the control blocks of all machines, 4 ints each
__global__ void k_group_ctl(const GroupDesc* __restrict__ gd, int count, int* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= count) return;
#pragma unroll
  for (int i = 0; i < 4; ++i) out[4 * q + i] = gd[q].ictl[i];
}
// x64 = x for the machines taking part (blockIdx.y = machine)
__global__ void k_group_finish(const GroupDesc* __restrict__ gd) {
  const GroupDesc& d = gd[blockIdx.y];
  if (!d.in) return;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < d.dim) d.x64[i] = (double)d.x[i];
}

// shape of the captured launches (a different shape needs a new capture)
void graph_key(const MixedWork& w, long key[11]) {
  const long k[11] = {w.ns, w.n,  w.m,  w.use_bias, w.Kp,         w.Np,
                      w.nsp, (long)w.dim, w.ct, w.chunk, w.chunk_iters};
  for (int i = 0; i < 11; ++i) key[i] = k[i];
}

// w.chunk_iters iterations (an even count: absolute iterations 1 + q chunk + p
// have parity (1 + p) & 1) captured once per operand shape, chunk length and
// preconditioner setting
int mixed_graph(MixedWork& w, const float* dinv, hipStream_t st, hipGraphExec_t* out,
                std::string& err) {
  long key[11];
  graph_key(w, key);
  bool same = true;
  for (int i = 0; i < 11; ++i) same = same && key[i] == w.gkey[i];
  if (!same) {
    mixed_drop_graphs(w);
    for (int i = 0; i < 11; ++i) w.gkey[i] = key[i];
  }
  const int slot = dinv ? 1 : 0;
  if (!w.gexec[slot]) {
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) return hipfail(err, "graph capture", e);
    for (int p = 0; p < w.chunk_iters; ++p) iteration_launches(w, dinv, (1 + p) & 1, false, st);
    publish_launch(w, st);
    e = hipStreamEndCapture(st, &g);
    if (e != hipSuccess) return hipfail(err, "graph capture", e);
    e = hipGraphInstantiate(&w.gexec[slot], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      w.gexec[slot] = nullptr;
      return hipfail(err, "graph instantiate", e);
    }
  }
  *out = w.gexec[slot];
  return YQ_OK;
}

// Tile-shape autotuner: times every (forward column tiles, backward chunk)
// pair on the prepared operands (3 applies after a warm-up, HIP events) and
// keeps the fastest per (Kp, Np, nsp) for the life of the process.  The
// shapes change only the order of float sums, not what is computed.
std::map<std::array<int, 3>, std::pair<int, int>> g_tuned;
std::mutex g_tuned_mu;

int tune(MixedWork& w, hipStream_t st, std::string& err) {
  const std::array<int, 3> key{w.Kp, w.Np, w.nsp};
  {
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) {
      w.ct = it->second.first;
      w.chunk = it->second.second;
      return YQ_OK;
    }
  }
  hipLaunchKernelGGL(k_zero_f, dim3(nb(w.dim)), dim3(256), 0, st, w.r, w.dim);
  hipLaunchKernelGGL(k_zero_f, dim3(nb(w.dim)), dim3(256), 0, st, w.p, w.dim);
  hipLaunchKernelGGL(k_zero_f, dim3(nb(w.dim)), dim3(256), 0, st, w.om32, w.dim);
  hipLaunchKernelGGL(k_set_ctl, dim3(1), dim3(64), 0, st, w.dctl, w.ictl, 1e-3, 0.0, 1);
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  if (e != hipSuccess) return hipfail(err, "autotune", e);
  float best = 1e30f;
  int bct = 4, bch = 256;
  for (int ct : {2, 4, 8}) {
    if (sizeof(uint16_t) * 3 * ct * 16 * (fwd_k_width(w.Kp) + kLdsPad) > 64 * 1024) continue;
    if (ct == 8 && w.Kp > 32 * kMaxKs) continue;  // the chunked forward takes 2 or 4
    for (int ch : {128, 256, 512}) {
      w.ct = ct;
      w.chunk = ch;
      w.ncg = (w.Np / 16 + ct - 1) / ct;
      apply_launches(w, st);
      e = hipEventRecord(a, st);
      for (int rep = 0; rep < 3 && e == hipSuccess; ++rep) apply_launches(w, st);
      if (e == hipSuccess) e = hipEventRecord(b, st);
      if (e == hipSuccess) e = hipEventSynchronize(b);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
      if (e == hipSuccess) e = hipGetLastError();
      if (e != hipSuccess) break;
      if (ms < best) {
        best = ms;
        bct = ct;
        bch = ch;
      }
    }
    if (e != hipSuccess) break;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (e != hipSuccess) return hipfail(err, "autotune", e);
  w.ct = bct;
  w.chunk = bch;
  std::lock_guard<std::mutex> lk(g_tuned_mu);
  g_tuned[key] = {bct, bch};
  return YQ_OK;
}

}  // namespace

void mixed_drop_graphs(MixedWork& w) {
  for (auto& g : w.gexec)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

int mixed_prepare(MixedWork& w, const int8_t* sigma, const float* T32, int ns, int n, int m,
                  int use_bias, hipStream_t st, std::string& err) {
  w.ns = ns;
  w.n = n;
  w.m = m;
  w.use_bias = use_bias;
  w.Kp = rup(n + 1, 32);
  w.Kr = rup(n + 1, 16);
  w.Np = rup(m + 1, 16);
  w.nsp = rup(ns, kRowPad);
  if (w.ct == 0) w.ct = 4;
  if (w.chunk == 0) w.chunk = 256;
  w.ncg = (w.Np / 16 + w.ct - 1) / w.ct;
  w.dim = (size_t)n * m + (use_bias ? (size_t)n + m : 0);
  const size_t rows = (size_t)w.nsp;
  const size_t red_total = red_layout(w, kBwdChunkMin).total;  // the most partials
  if (rows > w.cap_rows || (size_t)w.Kp > w.cap_kp || (size_t)w.Np > w.cap_np ||
      w.dim > w.cap_dim || red_total > w.cap_red) {
    const size_t chunks = rows / kBwdChunkMin;
    const size_t ncg_max = (w.Np / 16 + kColTilesMin - 1) / kColTilesMin;
    int rc;
    if ((rc = alloc(w.sigA, rows * sig_row_bytes(w.Kp), err)) ||
        (rc = alloc(w.sigT, rows / 8 * w.Kp, err)) ||
        (rc = alloc(w.Trm, rows * w.Np, err)) || (rc = alloc(w.Tt, rows * w.Np, err)) ||
        (rc = alloc(w.ypart, rows * ncg_max, err)) ||
        (rc = alloc(w.part, chunks * w.Kp * w.Np, err)) ||
        (rc = alloc(w.vsplit, (size_t)3 * w.Np * w.Kp, err)) ||
        (rc = alloc(w.om32, w.dim, err)) || (rc = alloc(w.x, w.dim, err)) ||
        (rc = alloc(w.r, w.dim, err)) || (rc = alloc(w.p, w.dim, err)) ||
        (rc = alloc(w.s, w.dim, err)) || (rc = alloc(w.u, w.dim, err)) ||
        (rc = alloc(w.Ap, w.dim, err)) || (rc = alloc(w.red, red_total, err)) ||
        (rc = alloc(w.dinv, w.dim, err)) || (rc = alloc(w.dctl, kCtlDoubles, err)) ||
        (rc = alloc(w.ictl, kCtlInts, err)))
      return rc;
    if (!w.host_ctl) {
      hipError_t e = hipHostMalloc((void**)&w.host_ctl, sizeof(int) * 2 * kSlotInts,
                                   hipHostMallocDefault);
      if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&w.host_ctl_dev, w.host_ctl, 0);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&w.ev[0], hipEventDisableTiming);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&w.ev[1], hipEventDisableTiming);
      if (e != hipSuccess) return hipfail(err, "mixed_prepare", e);
    }
    mixed_drop_graphs(w);  // buffers moved: captured launches are stale
    w.cap_rows = rows;
    w.cap_kp = w.Kp;
    w.cap_np = w.Np;
    w.cap_dim = w.dim;
    w.cap_red = red_total;
  }
  const int srow = sig_row_bytes(w.Kp);
  hipLaunchKernelGGL(k_prep_sig_rows, dim3(nb(rows * srow)), dim3(256), 0, st, sigma, ns, n,
                     w.nsp, srow, w.sigA);
  hipLaunchKernelGGL(k_prep_sig_cols, dim3(nb(rows / 8 * w.Kp)), dim3(256), 0, st, sigma, ns, n,
                     w.Kp, w.nsp, w.sigT);
  hipLaunchKernelGGL(k_prep_T, dim3(nb(rows * w.Np)), dim3(256), 0, st, T32, ns, m, w.Np, w.nsp,
                     w.Trm, w.Tt);
  // the CG step writes rows [0, Kr) of the split vector; the forward product
  // reads Kp rows, the rest zero
  {
    const hipError_t e = hipMemsetAsync(w.vsplit, 0, sizeof(uint16_t) * 3 * w.Np * w.Kp, st);
    if (e != hipSuccess) return hipfail(err, "mixed_prepare", e);
  }
  if (w.pin_ct && w.pin_chunk) {
    w.ct = w.pin_ct;
    w.chunk = w.pin_chunk;
  } else if (w.autotune) {
    const int rc = tune(w, st, err);
    if (rc) return rc;
  } else {
    w.ct = 4;
    w.chunk = 256;
  }
  if (w.ct == 8 && w.Kp > 32 * kMaxKs) {
    err = "the chunked forward product (n >= 160) takes 2 or 4 column tiles";
    return YQ_ERR_UNSUPPORTED;
  }
  if (sizeof(uint16_t) * 3 * w.ct * 16 * (fwd_k_width(w.Kp) + kLdsPad) > 64 * 1024) {
    err = "forward column group too wide for this n (LDS)";
    return YQ_ERR_UNSUPPORTED;
  }
  w.ncg = (w.Np / 16 + w.ct - 1) / w.ct;
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? YQ_OK : hipfail(err, "mixed_prepare", e);
}

int mixed_apply64(MixedWork& w, double shift, const double* omean64, const double* v,
                  double* out, hipStream_t st, std::string& err) {
  hipLaunchKernelGGL(k_set_ctl, dim3(1), dim3(64), 0, st, w.dctl, w.ictl, shift, 0.0, 1);
  hipLaunchKernelGGL(k_d2f, dim3(nb(w.dim)), dim3(256), 0, st, omean64, w.om32, w.dim);
  hipLaunchKernelGGL(k_d2f, dim3(nb(w.dim)), dim3(256), 0, st, v, w.r, w.dim);
  apply_launches(w, st);
  hipLaunchKernelGGL(k_f2d, dim3(nb(w.dim)), dim3(256), 0, st, w.Ap, out, w.dim);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? YQ_OK : hipfail(err, "mixed_apply", e);
}


int mixed_cg(MixedWork& w, double shift, const double* omean64, const double* diag64,
             const double* b, double* x, double tol, int max_iter, int* iters, hipStream_t st,
             std::string& err) {
  const size_t dim = w.dim;
  const RedLayout L = red_layout(w);
  hipLaunchKernelGGL(k_d2f, dim3(nb(dim)), dim3(256), 0, st, omean64, w.om32, dim);
  hipLaunchKernelGGL(k_d2f, dim3(nb(dim)), dim3(256), 0, st, b, w.r, dim);
  hipLaunchKernelGGL(k_zero_f, dim3(nb(dim)), dim3(256), 0, st, w.x, dim);
  const float* dinv = nullptr;
  if (diag64) {
    hipLaunchKernelGGL(k_jacobi_inv, dim3(nb(dim)), dim3(256), 0, st, diag64, shift, w.dinv, dim);
    dinv = w.dinv;
  }
  // start: u = M r with gamma, <r, r> and <u, u> in slot 0, and the products
  // on u (the control block is reset first so that the product kernels run);
  // then the threshold from <b, b>, formed on the device (b = 0 stops the run)
  hipLaunchKernelGGL(k_set_ctl, dim3(1), dim3(64), 0, st, w.dctl, w.ictl, shift, 0.0, 1);
  step_launch(w, dinv, 0, 0, st);
  product_launches(w, 0, st);
  hipLaunchKernelGGL(k_start_rr, dim3(1), dim3(256), 0, st, w.red + L.rr[0], L.nbv, tol,
                     max_iter, w.dctl, w.ictl);
  *iters = 0;
  // iteration 0 (no beta) eagerly, then chunks of w.chunk_iters iterations; an
  // iteration after the predicate fired leaves everything untouched, so the
  // chunk after the one that converged may already be queued
  iteration_launches(w, dinv, 0, true, st);
  hipError_t e;
  int rc;
  const bool graphs = w.use_graphs && !prof_enabled("cg_step") && !prof_enabled("fwd_mfma") &&
                      !prof_enabled("bwd_mfma");
  hipGraphExec_t gx = nullptr;
  if (graphs && (rc = mixed_graph(w, dinv, st, &gx, err))) return rc;
  // chunk c ends with k_publish (control block, <r, r> and the threshold into
  // host slot c & 1) and records event c & 1.  The host keeps one chunk
  // queued ahead of the one it polls, except when the <r, r> of the last two
  // polled chunks forecasts convergence inside the chunk already queued: a
  // chunk queued after the stop only runs kernels that return at once
  // (about 4.5 us each at N = 100, alpha = 4), a wrong forecast costs one
  // wait.  The forecast changes only when chunks are queued, never what the
  // iterations compute.
  int launched = 1, issued = 0;
  auto issue = [&]() -> hipError_t {
    if (gx) {
      hipError_t r = hipGraphLaunch(gx, st);
      if (r != hipSuccess) return r;
    } else {
      for (int k = 0; k < w.chunk_iters; ++k)
        iteration_launches(w, dinv, (launched + k) & 1, false, st);
      publish_launch(w, st);
    }
    launched += w.chunk_iters;
    hipError_t r = hipEventRecord(w.ev[issued & 1], st);
    ++issued;
    return r;
  };
  if (launched < max_iter && (e = issue()) != hipSuccess) return hipfail(err, "mixed_cg", e);
  if (issued == 0) {  // a budget of one iteration: publish the state after it
    publish_launch(w, st);
    if ((e = hipEventRecord(w.ev[0], st)) != hipSuccess) return hipfail(err, "mixed_cg", e);
    issued = 1;
  }
  // The state after the last polled chunk is final: a chunk is polled only
  // after every earlier one, the loop leaves at the stop flag (after which no
  // kernel changes the block) or after the last chunk issued.
  int last[4] = {0, 0, 0, 0};
  double lq_prev = 0.0, lq = 0.0;  // log(<r, r> / threshold) of the last two polls
  int it_prev = -1, it_last = -1;
  for (int k = 0; k < issued; ++k) {
    // iterations still needed after the last poll, at the rate of the last
    // two polls; the chunk being waited for covers w.chunk_iters of them
    bool ahead = launched < max_iter;
    if (ahead && k >= 2 && it_last > it_prev && lq < lq_prev && lq > 0.0) {
      const double need = lq * (double)(it_last - it_prev) / (lq_prev - lq);
      ahead = need > (double)w.chunk_iters;
    }
    if (ahead && (e = issue()) != hipSuccess) return hipfail(err, "mixed_cg", e);
    if ((e = hipEventSynchronize(w.ev[k & 1])) != hipSuccess) return hipfail(err, "mixed_cg", e);
    const int* c = w.host_ctl + kSlotInts * (k & 1);
    for (int i = 0; i < 4; ++i) last[i] = c[i];
    if (last[kCtlStop] || last[kCtlIters] >= max_iter) break;
    const double* cd = reinterpret_cast<const double*>(c + 4);
    lq_prev = lq;
    it_prev = it_last;
    lq = cd[0] > 0.0 && cd[1] > 0.0 ? log(cd[0] / cd[1]) : 0.0;
    it_last = last[kCtlIters];
    // not converged and nothing queued behind the polled chunk: queue one
    if (k + 1 == issued && launched < max_iter && (e = issue()) != hipSuccess)
      return hipfail(err, "mixed_cg", e);
  }
  hipLaunchKernelGGL(k_f2d, dim3(nb(dim)), dim3(256), 0, st, w.x, x, dim);
  e = hipGetLastError();
  if (e != hipSuccess) return hipfail(err, "mixed_cg", e);
  *iters = last[kCtlIters];
  // converged: the device's stop predicate saw <r, r> below the threshold
  // (or b = 0); a run that spent its budget with <r, r> above it reports
  // YQ_ERR_NOT_CONVERGED, a NaN or infinite <b, b> or <r, r> YQ_ERR_NAN
  if (last[kCtlConv] == kConvNonFinite) {
    err = "non-finite residual in the mixed-precision CG";
    return YQ_ERR_NAN;
  }
  return last[kCtlConv] == kConvYes ? YQ_OK : YQ_ERR_NOT_CONVERGED;
}

namespace {
// the three grouped launches of one CG product / step
void group_products(const MixedGroup& g, int count, hipStream_t st) {
  const GroupDesc* gd = static_cast<const GroupDesc*>(g.desc);
  const int2* bt = reinterpret_cast<const int2*>(g.btab);
  hipLaunchKernelGGL(k_fwd_group, dim3(g.nf), dim3(256), g.lds_f, st, gd, bt, count);
  hipLaunchKernelGGL(k_bwd_group, dim3(g.nbw), dim3(256), g.lds_b, st, gd, bt + g.nf);
}
void group_step(const MixedGroup& g, int mode, int par, hipStream_t st) {
  const GroupDesc* gd = static_cast<const GroupDesc*>(g.desc);
  const int2* bt = reinterpret_cast<const int2*>(g.btab);
  hipLaunchKernelGGL(k_cg_group, dim3(g.nc), dim3(g.cg_threads), 0, st, gd, bt + g.nf + g.nbw,
                     mode, par);
}
void group_iteration(const MixedGroup& g, int par, bool first, hipStream_t st) {
  group_step(g, first ? 1 : 2, par, st);
  group_products(g, 1, st);
}
}  // namespace

int mixed_cg_group(MixedGroup& g, std::vector<GroupJob>& jobs, double shift, hipStream_t st,
                   std::string& err) {
  const int count = (int)jobs.size();
  int maxit = 0;
  for (const GroupJob& j : jobs)
    if (j.active) maxit = std::max(maxit, j.max_iter);
  for (GroupJob& j : jobs) {
    j.iters = 0;
    j.rc = YQ_OK;
  }
  if (count == 0 || maxit <= 0) return YQ_OK;
  // descriptor and block tables (all machines, so the grids do not depend on
  // which machines iterate in this call)
  std::vector<GroupDesc> desc(count);
  std::vector<int2> bf, bb, bc;
  size_t lds_f = 0, lds_b = 0;
  long long maxdim = 0;
  int cgg = 1;
  for (int q = 0; q < count; ++q) {
    const GroupJob& j = jobs[q];
    MixedWork& w = *j.w;
    if (w.ncg != (w.Np / 16 + w.ct - 1) / w.ct || w.Kp > 32 * kMaxKs || w.Kp % 32 != 0) {
      err = "grouped CG: machine " + std::to_string(q) + " is not prepared";
      return YQ_ERR_STATE;
    }
    const RedLayout L = red_layout(w);
    GroupDesc d{};
    d.sigA = w.sigA;
    d.sigT = w.sigT;
    d.Trm = w.Trm;
    d.Tt = w.Tt;
    d.dinv = j.diag64 ? w.dinv : nullptr;
    d.om64 = j.omean64;
    d.b64 = j.b;
    d.diag64 = j.diag64;
    d.x64 = j.x;
    d.vs = w.vsplit;
    d.ypart = w.ypart;
    d.part = w.part;
    d.om = w.om32;
    d.x = w.x;
    d.r = w.r;
    d.p = w.p;
    d.s = w.s;
    d.u = w.u;
    d.Ap = w.Ap;
    d.dinv_w = w.dinv;
    d.red = w.red;
    d.dctl = w.dctl;
    d.ictl = w.ictl;
    d.dim = (long long)w.dim;
    for (int k = 0; k < 2; ++k) {
      d.g[k] = (long long)L.g[k];
      d.rr[k] = (long long)L.rr[k];
      d.uu[k] = (long long)L.uu[k];
    }
    d.op = (long long)L.op;
    d.ys = (long long)L.ys;
    d.yy = (long long)L.yy;
    d.tol = j.tol;
    d.ns = w.ns;
    d.n = w.n;
    d.m = w.m;
    d.use_bias = w.use_bias;
    d.Kp = w.Kp;
    d.Kr = w.Kr;
    d.Np = w.Np;
    d.nsp = w.nsp;
    d.ncg = w.ncg;
    d.ct = w.ct;
    d.chunk = w.chunk;
    d.nbv = L.nbv;
    d.nby = L.nby;
    d.fwd_gx = w.nsp / 256;
    d.bwd_gx = (w.Np / 16 + 3) / 4;
    d.bwd_total = d.bwd_gx * (w.nsp / w.chunk);
    d.cg_gx = w.Np / 16;
    d.cg_total = d.cg_gx * (w.Kr / 16);
    d.in = j.active ? 1 : 0;
    d.max_it = j.max_iter;
    desc[q] = d;
    for (int b = 0; b < d.fwd_gx * w.ncg; ++b) bf.push_back(make_int2(q, b));
    for (int b = 0; b < d.bwd_total; ++b) bb.push_back(make_int2(q, b));
    for (int b = 0; b < d.cg_total; ++b) bc.push_back(make_int2(q, b));
    lds_f = std::max(lds_f, sizeof(uint16_t) * 3 * w.ct * 16 * (w.Kp + kLdsPad));
    lds_b = std::max(lds_b, sizeof(uint16_t) * w.Kr * (kBwdSlice + kLdsPad));
    maxdim = std::max(maxdim, d.dim);
    cgg = std::max(cgg, cg_groups(w.nsp / w.chunk));
  }
  const int nblk = (int)(bf.size() + bb.size() + bc.size());
  hipError_t e = hipSuccess;
  if (count > g.cap || nblk > g.cap_blocks) {
    mixed_group_release(g);
    const int cap = std::max(count, 1), capb = std::max(nblk, 1);
    void* p = nullptr;
    if ((e = hipMalloc(&g.desc, sizeof(GroupDesc) * cap)) != hipSuccess ||
        (e = hipHostMalloc(&g.host_desc, sizeof(GroupDesc) * cap, hipHostMallocDefault)) !=
            hipSuccess ||
        (e = hipMalloc(&p, sizeof(int2) * capb)) != hipSuccess ||
        ((g.btab = static_cast<int*>(p)), false) ||
        (e = hipHostMalloc(&p, sizeof(int2) * capb, hipHostMallocDefault)) != hipSuccess ||
        ((g.host_btab = static_cast<int*>(p)), false) ||
        (e = hipMalloc(&p, sizeof(int) * 8 * cap)) != hipSuccess ||
        ((g.ctl = static_cast<int*>(p)), false) ||
        (e = hipHostMalloc(&p, sizeof(int) * 8 * cap, hipHostMallocDefault)) != hipSuccess ||
        ((g.host_ctl = static_cast<int*>(p)), false) ||
        (e = hipMalloc(&p, sizeof(double) * 3 * cap)) != hipSuccess ||
        ((g.stat = static_cast<double*>(p)), false) ||
        (e = hipHostMalloc(&p, sizeof(double) * 3 * cap, hipHostMallocDefault)) != hipSuccess ||
        ((g.host_stat = static_cast<double*>(p)), false) ||
        (e = hipEventCreateWithFlags(&g.ev[0], hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&g.ev[1], hipEventDisableTiming)) != hipSuccess) {
      mixed_group_release(g);
      return hipfail(err, "grouped CG allocation", e);
    }
    g.cap = cap;
    g.cap_blocks = capb;
  }
  // the staging buffers are free: the previous call ended with a synchronisation
  std::memcpy(g.host_desc, static_cast<const void*>(desc.data()), sizeof(GroupDesc) * count);
  int2* hb = reinterpret_cast<int2*>(g.host_btab);
  std::copy(bf.begin(), bf.end(), hb);
  std::copy(bb.begin(), bb.end(), hb + bf.size());
  std::copy(bc.begin(), bc.end(), hb + bf.size() + bb.size());
  const long key[7] = {count,       (long)bf.size(), (long)bb.size(), (long)bc.size(),
                       (long)lds_f, (long)lds_b,     256L * cgg};
  bool same = g.gexec != nullptr;
  for (int i = 0; i < 7; ++i) same = same && key[i] == g.gkey[i];
  if (!same && g.gexec) {
    (void)hipGraphExecDestroy(g.gexec);
    g.gexec = nullptr;
  }
  for (int i = 0; i < 7; ++i) g.gkey[i] = key[i];
  g.cg_threads = 256 * cgg;
  g.nf = (int)bf.size();
  g.nbw = (int)bb.size();
  g.nc = (int)bc.size();
  g.lds_f = lds_f;
  g.lds_b = lds_b;
  GroupDesc* gd = static_cast<GroupDesc*>(g.desc);
  if ((e = hipMemcpyAsync(g.desc, g.host_desc, sizeof(GroupDesc) * count,
                          hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(g.btab, g.host_btab, sizeof(int2) * nblk, hipMemcpyHostToDevice,
                          st)) != hipSuccess)
    return hipfail(err, "grouped CG", e);
  // start: u = M r and the products on u for every machine taking part, then
  // each machine's <b, b>, threshold and budget on the device
  hipLaunchKernelGGL(k_group_start, dim3(nb((size_t)maxdim), count), dim3(256), 0, st, gd,
                     shift);
  group_step(g, 0, 0, st);
  group_products(g, 0, st);
  hipLaunchKernelGGL(k_group_rr, dim3((count + 63) / 64), dim3(64), 0, st, gd, count, 0,
                     g.stat);
  // iteration 0 eagerly, then captured chunks of g.chunk_iters iterations
  // polled one chunk behind, as in mixed_cg
  group_iteration(g, 0, true, st);
  const bool graphs = g.use_graphs && !prof_enabled("cg_step") && !prof_enabled("fwd_mfma") &&
                      !prof_enabled("bwd_mfma");
  if (graphs && !g.gexec) {
    hipGraph_t gr = nullptr;
    e = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) return hipfail(err, "grouped CG capture", e);
    for (int p = 0; p < g.chunk_iters; ++p) group_iteration(g, (1 + p) & 1, false, st);
    e = hipStreamEndCapture(st, &gr);
    if (e != hipSuccess) return hipfail(err, "grouped CG capture", e);
    e = hipGraphInstantiate(&g.gexec, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) {
      g.gexec = nullptr;
      return hipfail(err, "grouped CG instantiate", e);
    }
  }
  int launched = 1, issued = 0;
  auto issue = [&]() -> hipError_t {
    if (graphs) {
      hipError_t r = hipGraphLaunch(g.gexec, st);
      if (r != hipSuccess) return r;
    } else {
      for (int k = 0; k < g.chunk_iters; ++k) group_iteration(g, (launched + k) & 1, false, st);
    }
    launched += g.chunk_iters;
    const int slot = issued & 1;
    hipLaunchKernelGGL(k_group_ctl, dim3((count + 63) / 64), dim3(64), 0, st, gd, count,
                       g.ctl + 4 * count * slot);
    hipError_t r = hipMemcpyAsync(g.host_ctl + 4 * count * slot, g.ctl + 4 * count * slot,
                                  sizeof(int) * 4 * count, hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipEventRecord(g.ev[slot], st);
    ++issued;
    return r;
  };
  if (launched < maxit && (e = issue()) != hipSuccess) return hipfail(err, "grouped CG", e);
  for (int k = 0; k < issued; ++k) {
    if (launched < maxit && (e = issue()) != hipSuccess) return hipfail(err, "grouped CG", e);
    if ((e = hipEventSynchronize(g.ev[k & 1])) != hipSuccess) return hipfail(err, "grouped CG", e);
    const int* c = g.host_ctl + 4 * count * (k & 1);
    bool all = true;
    for (int q = 0; q < count && all; ++q)
      all = c[4 * q + kCtlStop] || c[4 * q + kCtlIters] >= jobs[q].max_iter;
    if (all) break;
  }
  hipLaunchKernelGGL(k_group_rr, dim3((count + 63) / 64), dim3(64), 0, st, gd, count, 1,
                     g.stat);
  hipLaunchKernelGGL(k_group_finish, dim3(nb((size_t)maxdim), count), dim3(256), 0, st, gd);
  e = hipMemcpyAsync(g.host_stat, g.stat, sizeof(double) * 3 * count, hipMemcpyDeviceToHost,
                     st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess) e = hipGetLastError();
  if (e != hipSuccess) return hipfail(err, "grouped CG", e);
  for (int q = 0; q < count; ++q) {
    GroupJob& j = jobs[q];
    if (!j.active) continue;
    const double bn2 = g.host_stat[q];
    if (bn2 == 0.0) continue;  // x = 0 after no iterations
    if (!std::isfinite(bn2)) {
      j.rc = YQ_ERR_NAN;
      continue;
    }
    const double thr = j.tol * j.tol * bn2;
    const double rn2 = g.host_stat[count + q];
    j.iters = (int)g.host_stat[2 * count + q];
    j.rc = !std::isfinite(rn2) ? YQ_ERR_NAN : rn2 < thr ? YQ_OK : YQ_ERR_NOT_CONVERGED;
  }
  return YQ_OK;
}

void mixed_group_release(MixedGroup& g) {
  if (g.gexec) (void)hipGraphExecDestroy(g.gexec);
  for (void* p : {g.desc, (void*)g.btab, (void*)g.ctl, (void*)g.stat})
    if (p) (void)hipFree(p);
  for (void* p : {g.host_desc, (void*)g.host_btab, (void*)g.host_ctl, (void*)g.host_stat})
    if (p) (void)hipHostFree(p);
  for (auto& e : g.ev)
    if (e) (void)hipEventDestroy(e);
  const bool use_graphs = g.use_graphs;
  const int ci = g.chunk_iters;
  g = MixedGroup();
  g.use_graphs = use_graphs;
  g.chunk_iters = ci;
}

void mixed_release(MixedWork& w) {
  mixed_drop_graphs(w);
  for (void* p : {(void*)w.sigA, (void*)w.sigT, (void*)w.Trm, (void*)w.Tt, (void*)w.vsplit,
                  (void*)w.ypart, (void*)w.part, (void*)w.om32, (void*)w.x,
                  (void*)w.r, (void*)w.p, (void*)w.s, (void*)w.u, (void*)w.Ap, (void*)w.red,
                  (void*)w.dinv,
                  (void*)w.dctl, (void*)w.ictl})
    if (p) (void)hipFree(p);
  for (auto& e : w.ev)
    if (e) (void)hipEventDestroy(e);
  if (w.host_ctl) (void)hipHostFree(w.host_ctl);
  const bool g = w.use_graphs;
  w = MixedWork();
  w.use_graphs = g;
}

}  // namespace yq
