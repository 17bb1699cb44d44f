// S = alpha O_c^T O_c on FP64 MFMA (SRMat::corrMat, SRMat.hpp:142-149): the
// lower triangle of the dim x dim column-major S, one 128 x 128 tile per
// block, blocks over the lower block triangle.
//
// Operands.  Row s of the centred O is O_c[s][c] = sigma'_s,i(c) T'_s,j(c) -
// omean_c (scaled by rs_s for a weighted exact batch) with sigma' = [sigma |
// 1], T' = [tanh theta | 1] and (i(c), j(c)) the parameter's place on the
// (n + 1) x (m + 1) grid.  The factored form forms those values while it
// stages them in LDS, from the int8 spins and the FP64 T (Ns x m, 52 MB at
// N = 100, alpha = 4: cache resident, where O is 5.3 GB), so S never reads
// O; the dense form reads a stored row-major O (the FeedForward block of an
// AmplitudePhase machine).
//
// v_mfma_f64_16x16x4f64: lane l holds A[row l%16][k l/16] and B[k l/16][col
// l%16] and returns D[(l/16) + 4q][l%16], q = 0..3.  The tile is formed
// transposed, D = O_c[:, J]^T O_c[:, I] (rows j, columns i, J the smaller
// block), so that for each q the 16 lanes of a lane group hold 16
// consecutive i of one j: the store into S[i + j dim] (column-major, i >= j)
// is then 128 contiguous bytes per lane group.
//
// Block: 256 threads = 4 waves in 2 x 2, each a 64 x 64 sub-tile (4 x 4 MFMA
// tiles, 16 double4 accumulators).  K (the samples) advances in stages of 16
// samples: both 16 x 128 operand slabs are staged in LDS (double-buffered; the
// next stage's global loads are in flight during the current stage's 64
// MFMAs per wave).
#include <math.h>

#include <algorithm>

#include "yq_internal.h"

namespace yq {
namespace {

typedef double double4v __attribute__((ext_vector_type(4)));

constexpr int kT = 128;           // tile edge
constexpr int kBK = 16;           // samples per stage
constexpr int kLd = kT + 4;       // LDS row stride (doubles)
constexpr int kPer = kBK * kT / 256;  // slab elements per thread and stage (8)

struct SyrkArgs {
  // factored operands (dense == 0)
  const int8_t* sigma;   // ns x n
  const double* T;       // ns x m, tanh theta
  const double* omean;   // dim
  const double* rs;      // ns row scales, or null
  int n, m, use_bias;
  // dense operand (dense == 1): row-major ns x dim
  const double* O;
  int dense;
  int ns;
  long long dim;
  double alpha;
  double* S;             // dim x dim column-major, lower triangle written
  int nb;                // tiles per edge
  int kchunk;            // samples per blockIdx.y chunk (the whole batch when one)
  size_t pstride;        // doubles between the chunks' partial products
};

// (bi, bj), bi >= bj, of linear tile t of the lower block triangle (row-major
// over bi): t = bi (bi + 1) / 2 + bj
__device__ inline void tri_tile(int t, int& bi, int& bj) {
  int r = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((long long)r * (r + 1) / 2 > t) --r;
  while ((long long)(r + 1) * (r + 2) / 2 <= t) ++r;
  bi = r;
  bj = t - r * (r + 1) / 2;
}

// grid place of parameter c: W block c = i m + j -> (i, j); a (visible bias)
// c = nm + i -> (i, m); b (hidden bias) c = nm + n + j -> (n, j)
__device__ inline void grid_of(long long c, int n, int m, int& i, int& j) {
  const long long nm = (long long)n * m;
  if (c < nm) {
    i = (int)(c / m);
    j = (int)(c - (long long)i * m);
  } else if (c < nm + n) {
    i = (int)(c - nm);
    j = m;
  } else {
    i = n;
    j = (int)(c - nm - n);
  }
}

// One staged column of a slab.  The loads of a stage are issued without
// branches -- every address is a valid one (a column past dim reads column
// dim - 1, a sample past the chunk the chunk's last sample, the bias rows and
// columns of the augmented grid read spin 0 / T column 0) -- and the values
// are formed when the stage is stored to LDS, after the previous stage's
// MFMAs, so that all of a stage's loads are in flight together.
// Factored: tp = T + j, sp = spins + i; use_t = 0 for the a-bias column
// (T' = 1), use_s = 0 for the b-bias row (sigma' = 1); om = omean_c.  Dense:
// tp = O + c.  live = 0 past dim (the value is 0).
struct Col {
  const int8_t* sp;
  const double* tp;
  double om;
  int use_s, use_t, live;
};

template <bool DENSE>
__device__ inline Col make_col(const SyrkArgs& a, long long c) {
  Col q;
  q.live = c < a.dim;
  const long long cc = q.live ? c : a.dim - 1;
  if (DENSE) {
    q.tp = a.O + cc;
    q.sp = nullptr;
    q.om = 0.0;
    q.use_s = 0;
    q.use_t = 1;
  } else {
    int i, j;
    grid_of(cc, a.n, a.m, i, j);
    q.use_s = i < a.n;
    q.use_t = j < a.m;
    q.sp = a.sigma + (q.use_s ? i : 0);
    q.tp = a.T + (q.use_t ? j : 0);
    q.om = a.omean[cc];
  }
  return q;
}

template <bool DENSE, bool RS>
__global__ void __launch_bounds__(256, 2) k_syrk64_mfma(const SyrkArgs a) {
  __shared__ double sh[2][2][kBK][kLd];  // [buffer][J slab, I slab][sample][column]
  int bi, bj;
  tri_tile(blockIdx.x, bi, bj);
  const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
  const int li = l & 15, lg = l >> 4;
  const int wr = wave >> 1, wc = wave & 1;  // rows (J) and columns (I) of the sub-tile
  // the samples of this block's chunk: [s0, s1)
  const int s0 = blockIdx.y * a.kchunk;
  const int s1 = min(a.ns, s0 + a.kchunk);
  // the staged column of this thread in both slabs, and its samples
  const int cl = tid & (kT - 1), s_off = tid >> 7;  // samples s_off + 2 p, p < kPer
  const Col cj = make_col<DENSE>(a, (long long)bj * kT + cl);
  const Col ci = make_col<DENSE>(a, (long long)bi * kT + cl);
  const size_t tstride = DENSE ? (size_t)a.dim : (size_t)a.m;
  // raw loads of one stage: T (or O) values, spins and row scales
  double tj[kPer], ti[kPer], rsv[kPer];
  int8_t gj[kPer], gi[kPer];
  int k_loaded = s0;
  auto load = [&](int k0) {
    k_loaded = k0;
#pragma unroll
    for (int p = 0; p < kPer; ++p) {
      const int s = min(k0 + s_off + 2 * p, s1 - 1);
      tj[p] = cj.tp[(size_t)s * tstride];
      ti[p] = ci.tp[(size_t)s * tstride];
      if (!DENSE) {
        gj[p] = cj.sp[(size_t)s * a.n];
        gi[p] = ci.sp[(size_t)s * a.n];
      }
      if (RS) rsv[p] = a.rs[s];
    }
  };
  auto value = [&](const Col& q, double t, int8_t g, double r, bool in) {
    double v = t;
    if (!DENSE) {
      v = q.use_t ? t : 1.0;
      v = (q.use_s && g < 0) ? -v : v;
      v = v - q.om;
    }
    if (RS) v = v * r;
    return in && q.live ? v : 0.0;
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int p = 0; p < kPer; ++p) {
      const bool in = k_loaded + s_off + 2 * p < s1;
      sh[buf][0][s_off + 2 * p][cl] = value(cj, tj[p], DENSE ? 0 : gj[p], RS ? rsv[p] : 1.0, in);
      sh[buf][1][s_off + 2 * p][cl] = value(ci, ti[p], DENSE ? 0 : gi[p], RS ? rsv[p] : 1.0, in);
    }
  };
  double4v acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = (double4v){0.0, 0.0, 0.0, 0.0};
  const int stages = (s1 - s0 + kBK - 1) / kBK;
  load(s0);
  store(0);
  __syncthreads();
  for (int st = 0; st < stages; ++st) {
    const int buf = st & 1;
    if (st + 1 < stages) load(s0 + (st + 1) * kBK);
#pragma unroll
    for (int ks = 0; ks < kBK / 4; ++ks) {
      double av[4], bv[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) av[r] = sh[buf][0][ks * 4 + lg][wr * 64 + r * 16 + li];
#pragma unroll
      for (int c = 0; c < 4; ++c) bv[c] = sh[buf][1][ks * 4 + lg][wc * 64 + c * 16 + li];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c)
          acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[r], bv[c], acc[r][c], 0, 0, 0);
    }
    if (st + 1 < stages) store(buf ^ 1);
    __syncthreads();
  }
  // D[r][c] = sum_s O_c[s][j_r] O_c[s][i_c] = S[i_c][j_r]: lane l holds rows
  // j = .. + lg + 4q and column i = .. + li; for fixed q the 16 lanes of a
  // group store 16 consecutive i of column j of the column-major S
  const long long dim = a.dim;
  double* out = a.S + (size_t)blockIdx.y * a.pstride;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const long long i = (long long)bi * kT + wc * 64 + c * 16 + li;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const long long j = (long long)bj * kT + wr * 64 + r * 16 + lg + 4 * q;
        if (i < dim && j <= i) out[(size_t)j * dim + i] = a.alpha * acc[r][c][q];
      }
    }
}

// ---- split-bf16 S (mixed precision) ---------------------------------------
// O_c values rounded to float and split x = h + m + l into three bf16 (24
// significant bits); each product keeps the six pair terms down to 2^-24 of
// the leading one, hh + hm + mh + hl + lh + mm, on v_mfma_f32_16x16x32_bf16
// with fp32 accumulation (lane l holds A[row l%16][k 8(l/16) .. +7] and
// B[k 8(l/16) .. +7][col l%16] and returns D[4(l/16) + q][l%16]); S is
// stored in double.  Tile 128 x 128 as above, 32 samples per stage, the
// three layers of both slabs in LDS as [layer][column][sample] so that a
// lane's eight consecutive samples are one 16-byte read.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
constexpr int kBK3 = 32;             // samples per stage
constexpr int kLd3 = kBK3 + 8;       // LDS row stride (bf16)
constexpr int kPer3 = kBK3 * kT / 256;  // slab elements per thread and stage (16)

__device__ inline uint16_t to_bf16(float x) {
  uint32_t u = __builtin_bit_cast(uint32_t, x);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ inline float from_bf16(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

template <bool DENSE, bool RS>
__global__ void __launch_bounds__(256, 2) k_syrk_bf16x3(const SyrkArgs a) {
  __shared__ uint16_t sh[2][3][kT][kLd3];  // [J, I slab][hi, mid, lo][column][sample]
  int bi, bj;
  tri_tile(blockIdx.x, bi, bj);
  const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
  const int li = l & 15, lg = l >> 4;
  const int wr = wave >> 1, wc = wave & 1;
  const int s0 = blockIdx.y * a.kchunk;
  const int s1 = min(a.ns, s0 + a.kchunk);
  // this thread stages column cl of both slabs for 16 consecutive samples
  const int cl = tid & (kT - 1), sb = (tid >> 7) * kPer3;
  const Col cj = make_col<DENSE>(a, (long long)bj * kT + cl);
  const Col ci = make_col<DENSE>(a, (long long)bi * kT + cl);
  const size_t tstride = DENSE ? (size_t)a.dim : (size_t)a.m;
  double tj[kPer3], ti[kPer3], rsv[kPer3];
  int8_t gj[kPer3], gi[kPer3];
  int k_loaded = s0;
  auto load = [&](int k0) {
    k_loaded = k0;
#pragma unroll
    for (int p = 0; p < kPer3; ++p) {
      const int s = min(k0 + sb + p, s1 - 1);
      tj[p] = cj.tp[(size_t)s * tstride];
      ti[p] = ci.tp[(size_t)s * tstride];
      if (!DENSE) {
        gj[p] = cj.sp[(size_t)s * a.n];
        gi[p] = ci.sp[(size_t)s * a.n];
      }
      if (RS) rsv[p] = a.rs[s];
    }
  };
  auto value = [&](const Col& q, double t, int8_t g, double r, bool in) {
    double v = t;
    if (!DENSE) {
      v = q.use_t ? t : 1.0;
      v = (q.use_s && g < 0) ? -v : v;
      v = v - q.om;
    }
    if (RS) v = v * r;
    return in && q.live ? (float)v : 0.0f;
  };
  auto store = [&]() {
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      u16x8 hj[3], hi_[3];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int p = half * 8 + e;
        const bool in = k_loaded + sb + p < s1;
        const float x = value(cj, tj[p], DENSE ? 0 : gj[p], RS ? rsv[p] : 1.0, in);
        const float y = value(ci, ti[p], DENSE ? 0 : gi[p], RS ? rsv[p] : 1.0, in);
        const uint16_t xh = to_bf16(x), yh = to_bf16(y);
        const float xr = x - from_bf16(xh), yr = y - from_bf16(yh);
        const uint16_t xm = to_bf16(xr), ym = to_bf16(yr);
        hj[0][e] = xh;
        hj[1][e] = xm;
        hj[2][e] = to_bf16(xr - from_bf16(xm));
        hi_[0][e] = yh;
        hi_[1][e] = ym;
        hi_[2][e] = to_bf16(yr - from_bf16(ym));
      }
#pragma unroll
      for (int ly = 0; ly < 3; ++ly) {
        *reinterpret_cast<u16x8*>(&sh[0][ly][cl][sb + half * 8]) = hj[ly];
        *reinterpret_cast<u16x8*>(&sh[1][ly][cl][sb + half * 8]) = hi_[ly];
      }
    }
  };
  floatx4 acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = (floatx4){0.f, 0.f, 0.f, 0.f};
  const int stages = (s1 - s0 + kBK3 - 1) / kBK3;
  load(s0);
  for (int st = 0; st < stages; ++st) {
    __syncthreads();  // the previous stage's MFMAs are done with the slabs
    store();
    __syncthreads();
    if (st + 1 < stages) load(s0 + (st + 1) * kBK3);
    bf16x8 av[4][3];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int ly = 0; ly < 3; ++ly)
        av[r][ly] = __builtin_bit_cast(
            bf16x8, *reinterpret_cast<const u16x8*>(&sh[0][ly][wr * 64 + r * 16 + li][lg * 8]));
    // each stage's 32 samples start from zero, smallest pair products first,
    // and join the running sum through a round-to-nearest add: chaining the
    // MFMA accumulator over every stage lets its rounding build up with the
    // sample count (1e-5 of max|S| at 5,000 samples)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      bf16x8 bv[3];
#pragma unroll
      for (int ly = 0; ly < 3; ++ly)
        bv[ly] = __builtin_bit_cast(
            bf16x8, *reinterpret_cast<const u16x8*>(&sh[1][ly][wc * 64 + c * 16 + li][lg * 8]));
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        floatx4 t = (floatx4){0.f, 0.f, 0.f, 0.f};
        t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[r][0], bv[2], t, 0, 0, 0);
        t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[r][2], bv[0], t, 0, 0, 0);
        t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[r][1], bv[1], t, 0, 0, 0);
        t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[r][0], bv[1], t, 0, 0, 0);
        t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[r][1], bv[0], t, 0, 0, 0);
        t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[r][0], bv[0], t, 0, 0, 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[r][c][q] = acc[r][c][q] + t[q];
      }
    }
  }
  // D[4 lg + q][li] = sum_s O_c[s][j] O_c[s][i] with j the row (A = J slab):
  // for fixed q the 16 lanes of a group hold 16 consecutive i of one j
  const long long dim = a.dim;
  double* out = a.S + (size_t)blockIdx.y * a.pstride;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const long long i = (long long)bi * kT + wc * 64 + c * 16 + li;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const long long j = (long long)bj * kT + wr * 64 + r * 16 + 4 * lg + q;
        if (i < dim && j <= i) out[(size_t)j * dim + i] = a.alpha * (double)acc[r][c][q];
      }
    }
}

// A complex S = A A^H (A = conj(O_c), the stored rows) from the real product
// XX = X^T X of X = the same rows read as 2 dim reals (re, im interleaved):
// with A = R + i I, S_ij = sum_s (R_si + i I_si)(R_sj - i I_sj), so
// Re S_ij = XX(2i, 2j) + XX(2i+1, 2j+1) and Im S_ij = XX(2i+1, 2j) -
// XX(2i, 2j+1), both in XX's lower triangle for i > j (Im S_ii = 0).  XX and S
// are column-major lower triangles; S is complex interleaved.
__global__ void k_herk_combine(const double* __restrict__ XX, size_t dim, double* __restrict__ S) {
  const size_t total = dim * dim, d2 = 2 * dim;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (size_t)gridDim.x * blockDim.x) {
    const size_t i = q % dim, j = q / dim;
    if (i < j) continue;
    const double re = XX[2 * i + 2 * j * d2] + XX[(2 * i + 1) + (2 * j + 1) * d2];
    const double im = i == j ? 0.0 : XX[(2 * i + 1) + 2 * j * d2] - XX[2 * i + (2 * j + 1) * d2];
    S[2 * q] = re;
    S[2 * q + 1] = im;
  }
}

SyrkArgs base_args(int ns, size_t dim, double alpha, double* S, int chunks, double* part) {
  SyrkArgs a{};
  a.ns = ns;
  a.dim = (long long)dim;
  a.alpha = alpha;
  a.nb = (int)((dim + kT - 1) / kT);
  a.kchunk = chunks > 1 ? (ns + chunks - 1) / chunks : ns;
  a.S = chunks > 1 ? part : S;
  a.pstride = dim * dim;
  return a;
}

void launch(const SyrkArgs& a, int chunks, double* S, hipStream_t st) {
  const long long tiles = (long long)a.nb * (a.nb + 1) / 2;
  ProfGuard g("syrk64", st);
  const dim3 grid((unsigned)tiles, (unsigned)chunks), block(256);
  if (a.dense)
    hipLaunchKernelGGL((k_syrk64_mfma<true, false>), grid, block, 0, st, a);
  else if (a.rs)
    hipLaunchKernelGGL((k_syrk64_mfma<false, true>), grid, block, 0, st, a);
  else
    hipLaunchKernelGGL((k_syrk64_mfma<false, false>), grid, block, 0, st, a);
  // a split batch: the chunks' partial products added in chunk order
  if (chunks > 1) launch_sum_lower(a.S, chunks, (size_t)a.dim, 1, S, st);
}

}  // namespace

int syrk64_chunks(int ns, size_t dim, size_t part_cap_doubles) {
  const long long nb = (long long)((dim + kT - 1) / kT);
  const long long tiles = nb * (nb + 1) / 2;
  if (tiles >= 1024) return 1;
  long long P = (1024 + tiles - 1) / tiles;
  P = std::min<long long>(P, ns / 4096);                       // >= 4096 samples a chunk
  P = std::min<long long>(P, (long long)(part_cap_doubles / (dim * dim)));
  return P > 1 ? (int)P : 1;
}

int launch_syrk64_factored(const int8_t* sigma, const double* T, const double* omean,
                           const double* rs, int ns, int n, int m, int use_bias, double alpha,
                           double* S, int chunks, double* part, hipStream_t st) {
  const size_t dim = use_bias ? (size_t)n * m + n + m : (size_t)n * m;
  if (ns < 1 || dim < 1 || chunks < 1 || (chunks > 1 && !part)) return YQ_ERR_ARG;
  SyrkArgs a = base_args(ns, dim, alpha, S, chunks, part);
  a.sigma = sigma;
  a.T = T;
  a.omean = omean;
  a.rs = rs;
  a.n = n;
  a.m = m;
  a.use_bias = use_bias;
  launch(a, chunks, S, st);
  return YQ_OK;
}

int launch_syrk_bf16x3(const int8_t* sigma, const double* T, const double* omean,
                       const double* rs, int ns, int n, int m, int use_bias, double alpha,
                       double* S, int chunks, double* part, hipStream_t st) {
  const size_t dim = use_bias ? (size_t)n * m + n + m : (size_t)n * m;
  if (ns < 1 || dim < 1 || chunks < 1 || (chunks > 1 && !part)) return YQ_ERR_ARG;
  if (rs) return YQ_ERR_UNSUPPORTED;  // weighted (exact) batches use the FP64 SYRK
  SyrkArgs a = base_args(ns, dim, alpha, S, chunks, part);
  a.sigma = sigma;
  a.T = T;
  a.omean = omean;
  a.n = n;
  a.m = m;
  a.use_bias = use_bias;
  const long long tiles = (long long)a.nb * (a.nb + 1) / 2;
  {
    ProfGuard g("syrk_bf16x3", st);
    const dim3 grid((unsigned)tiles, (unsigned)chunks), block(256);
    hipLaunchKernelGGL((k_syrk_bf16x3<false, false>), grid, block, 0, st, a);
  }
  if (chunks > 1) launch_sum_lower(a.S, chunks, dim, 1, S, st);
  return YQ_OK;
}

int launch_syrk_bf16x3_dense(const double* O, int ns, size_t dim, double alpha, double* S,
                             int chunks, double* part, hipStream_t st) {
  if (ns < 1 || dim < 1 || chunks < 1 || (chunks > 1 && !part)) return YQ_ERR_ARG;
  SyrkArgs a = base_args(ns, dim, alpha, S, chunks, part);
  a.O = O;
  a.dense = 1;
  const long long tiles = (long long)a.nb * (a.nb + 1) / 2;
  {
    ProfGuard g("syrk_bf16x3", st);
    const dim3 grid((unsigned)tiles, (unsigned)chunks), block(256);
    hipLaunchKernelGGL((k_syrk_bf16x3<true, false>), grid, block, 0, st, a);
  }
  if (chunks > 1) launch_sum_lower(a.S, chunks, dim, 1, S, st);
  return YQ_OK;
}

void launch_herk_combine(const double* XX, size_t dim, double* S, hipStream_t st) {
  const size_t total = dim * dim;
  const unsigned grid = (unsigned)std::min<size_t>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(k_herk_combine, dim3(grid), dim3(256), 0, st, XX, dim, S);
}

int launch_syrk64_dense(const double* O, int ns, size_t dim, double alpha, double* S, int chunks,
                        double* part, hipStream_t st) {
  if (ns < 1 || dim < 1 || chunks < 1 || (chunks > 1 && !part)) return YQ_ERR_ARG;
  SyrkArgs a = base_args(ns, dim, alpha, S, chunks, part);
  a.O = O;
  a.dense = 1;
  launch(a, chunks, S, st);
  return YQ_OK;
}

}  // namespace yq
