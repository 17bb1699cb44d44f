// General FeedForward stack on gfx950: any sequence of Conv1D (channels,
// stride, bias), FullyConnected (bias), activation and Mean layers, as
// FeedForward::forward / makeData / backward run it (Machines/
// FeedForward.hpp:152-205).
//
// One 256-thread block owns sb samples at a time.  Every activation of a
// sample -- the input spins and each layer's output, the vectors of makeData
// -- sits in one LDS strip, and the backward sweep's dL/d(out) vectors in two
// ping-pong LDS buffers of the widest layer, so a layer is one pass of the
// block over (sample, output) pairs followed by a barrier.  The layer table
// lives in device memory and is read with uniform (scalar) loads.  Modes as
// the C4 kernel (ff.hip): log psi (optionally the strips), the Jacobian rows
// (each thread writes the derivatives it forms, parameter index fastest, so
// a FullyConnected block of a row goes out as coalesced stores), and the VJP
// grad = sum_s r_s O_s through per-block partial rows (each entry owned by one
// thread of the block: the block's first sample group writes it, later
// groups add, in group order) and the fixed-order block reduction.
#include "yq_internal.h"
#include "yq_math.h"

namespace yq {
namespace {

enum { kFwd = 0, kJac = 1, kVjp = 2 };

// activations.hpp: value a = f(z) and G = J^T F from (z, a, F)
template <typename R>
__device__ inline R act_fwd(int act, R prm, R z) {
  switch (act) {
    case YQ_ACT_LNCOSH:  // :85-89 logCosh
      return Math<R>::logcosh(z);
    case YQ_ACT_TANH:  // :126-128 a tanh(z / a)
      return prm * tanh(z / prm);
    case YQ_ACT_WEAKTANH:  // :168-171 z - tanh(z) / 2
      return z - tanh(z) / (R)2;
    case YQ_ACT_SIGMOID:  // :214-219: exp(z), then += 1 (the two inverse()
                          // calls of the reference are discarded expressions)
      return exp(z) + (R)1;
    case YQ_ACT_RELU:  // :303-307
      return z > (R)0 ? z : (R)0;
    case YQ_ACT_LEAKYRELU:  // :404-408
      return z > (R)0 ? z : prm * z;
    case YQ_ACT_HARDTANH:  // :451-458
      return fabs(z) > (R)1 ? copysign((R)1, z) : z;
    case YQ_ACT_SOFTSHRINK:  // :500-507
      return fabs(z) < (R)1 ? (R)0 : z - copysign((R)1, z);
    case YQ_ACT_LEAKYHARDTANH: {  // :553-562
      R v = fabs(z) > (R)1 ? copysign((R)1, z) : z;
      v += fabs(z) < (R)1 ? (R)0 : prm * (z - copysign((R)1, z));
      return v;
    }
    case YQ_ACT_SOFTSIGN:  // :605-607
      return z / (fabs(z) + (R)1);
    case YQ_ACT_COS:  // :651-653 a cos(pi z / a)
      return prm * cos(z * (R)M_PI / prm);
    default:  // Identity
      return z;
  }
}
template <typename R>
__device__ inline R act_bwd(int act, R prm, R z, R a, R F) {
  switch (act) {
    case YQ_ACT_LNCOSH:  // :96-100 F tanh(z)
      return F * tanh(z);
    case YQ_ACT_TANH:  // :137-141 F (1 - a^2 / c^2)
      return F * ((R)1 - a * a / (prm * prm));
    case YQ_ACT_WEAKTANH: {  // :179-184 F (1 + tanh^2) / 2
      const R t = tanh(z);
      return F * ((R)1 + t * t) / (R)2;
    }
    case YQ_ACT_SIGMOID:  // :226-230 (1 - a) a, F unused by the reference
      return ((R)1 - a) * a;
    case YQ_ACT_RELU:  // :314-320
      return z > (R)0 ? F : (R)0;
    case YQ_ACT_LEAKYRELU:  // :415-421
      return z > (R)0 ? F : prm * F;
    case YQ_ACT_HARDTANH:  // :465-471
      return fabs(z) > (R)1 ? (R)0 : F;
    case YQ_ACT_SOFTSHRINK:  // :514-520
      return fabs(z) < (R)1 ? (R)0 : F;
    case YQ_ACT_LEAKYHARDTANH:  // :569-575
      return fabs(z) > (R)1 ? prm * F : F;
    case YQ_ACT_SOFTSIGN: {  // :614-617 F / (1 + |z|)^2
      const R d = (R)1 + fabs(z);
      return ((R)1 / (d * d)) * F;
    }
    case YQ_ACT_COS:  // :660-663 -pi sin(pi z / a) F
      return -(R)M_PI * sin((R)M_PI * z / prm) * F;
    default:
      return F;
  }
}

// Conv1D input site of output r, kernel entry ki (layers/Conv1D.hpp:160-162)
__device__ inline int conv_site(int r, int ki, int stride, int K, int inSize) {
  return (r * stride + ki - K / 2 + inSize) % inSize;
}

// Forward of layer L for cnt samples: strip s holds the input at in_off and
// receives the output at out_off.
template <typename R>
__device__ void layer_forward(const FFLayerDev& L, const R* __restrict__ P, R* act, int AL,
                              int cnt) {
  const int tid = threadIdx.x;
  const int od = L.out_dim, id = L.in_dim;
  const R* w = P + L.p_off;
  for (int q = tid; q < cnt * od; q += 256) {
    const int s = q / od, o = q % od;
    const R* in = act + (size_t)s * AL + L.in_off;
    R v;
    if (L.kind == YQ_LAYER_CONV1D) {
      const int inSize = id / L.ic, outSize = od / L.oc;
      const int oc = o / outSize, r = o % outSize;
      const R* kc = w + (size_t)oc * L.ic * L.K;  // column oc of the kernel matrix
      v = (R)0;
      for (int ic = 0; ic < L.ic; ++ic)
        for (int ki = 0; ki < L.K; ++ki)
          v += kc[ki + ic * L.K] * in[conv_site(r, ki, L.stride, L.K, inSize) + ic * inSize];
      if (L.bias) v += w[(size_t)L.ic * L.K * L.oc + oc];
    } else if (L.kind == YQ_LAYER_FC) {
      const R* col = w + (size_t)o * id;  // W(:, o), column-major in x out
      R t = (R)0;
      for (int i = 0; i < id; ++i) t += col[i] * in[i];
      v = L.bias ? w[(size_t)id * od + o] + t : t;
    } else if (L.kind == YQ_LAYER_ACT) {
      v = act_fwd<R>(L.act, (R)L.prm, in[o]);
    } else {  // Mean: input rows x cols column-major
      R t = (R)0;
      if (L.axis == 0) {
        for (int i = 0; i < L.rows; ++i) t += in[i + o * L.rows];
        v = t / (R)L.rows;
      } else {
        for (int j = 0; j < L.cols; ++j) t += in[o + j * L.rows];
        v = t / (R)L.cols;
      }
    }
    act[(size_t)s * AL + L.out_off + o] = v;
  }
}

// d(out)/d(parameter c) contracted with dout for one sample of layer L
// (Conv1D.hpp:196-220, FullyConnected.hpp:168-172)
template <typename R>
__device__ inline R layer_der(const FFLayerDev& L, const R* in, const R* dout, int c) {
  if (L.kind == YQ_LAYER_CONV1D) {
    const int inSize = L.in_dim / L.ic, outSize = L.out_dim / L.oc;
    const int nk = L.ic * L.K * L.oc;
    R v = (R)0;
    if (c < nk) {
      const int ki = c % L.K, ic = (c / L.K) % L.ic, oc = c / (L.K * L.ic);
      for (int r = 0; r < outSize; ++r)
        v += in[conv_site(r, ki, L.stride, L.K, inSize) + ic * inSize] * dout[r + oc * outSize];
    } else {
      const int oc = c - nk;
      for (int r = 0; r < outSize; ++r) v += dout[r + oc * outSize];
    }
    return v;
  }
  // FullyConnected: W(i, o) at i + o * in, then the bias
  const int nw = L.in_dim * L.out_dim;
  if (c < nw) return in[c % L.in_dim] * dout[c / L.in_dim];
  return dout[c - nw];
}

// dL/d(in) of layer L (Conv1D.hpp:185-191, FullyConnected.hpp:175-176,
// ActivationLayer.hpp:108-114, Mean.hpp:112-123)
template <typename R>
__device__ void layer_din(const FFLayerDev& L, const R* __restrict__ P, const R* act, int AL,
                          const R* dcur, R* dnext, int D, int cnt) {
  const int tid = threadIdx.x;
  const int id = L.in_dim, od = L.out_dim;
  const R* w = P + L.p_off;
  for (int q = tid; q < cnt * id; q += 256) {
    const int s = q / id, p = q % id;
    const R* dout = dcur + (size_t)s * D;
    R v = (R)0;
    if (L.kind == YQ_LAYER_CONV1D) {
      // every (oc, r, ki) whose periodic site is p: r * stride =
      // p + j inSize + K/2 - ki - inSize for j = 0, 1, 2
      const int inSize = id / L.ic, outSize = od / L.oc;
      const int ic = p / inSize, pp = p % inSize;
      for (int oc = 0; oc < L.oc; ++oc) {
        const R* kc = w + (size_t)oc * L.ic * L.K + ic * L.K;
        for (int ki = 0; ki < L.K; ++ki)
          for (int j = 0; j < 3; ++j) {
            const int rs = pp + (j - 1) * inSize + L.K / 2 - ki;
            if (rs < 0 || rs % L.stride) continue;
            const int r = rs / L.stride;
            if (r < outSize) v += kc[ki] * dout[r + oc * outSize];
          }
      }
    } else if (L.kind == YQ_LAYER_FC) {
      for (int o = 0; o < od; ++o) v += w[p + (size_t)o * id] * dout[o];
    } else if (L.kind == YQ_LAYER_ACT) {
      const R* st = act + (size_t)s * AL;
      v = act_bwd<R>(L.act, (R)L.prm, st[L.in_off + p], st[L.out_off + p], dout[p]);
    } else {
      v = L.axis == 0 ? dout[p / L.rows] / (R)L.rows : dout[p % L.rows] / (R)L.cols;
    }
    dnext[(size_t)s * D + p] = v;
  }
}

template <typename R, int MODE>
__global__ void __launch_bounds__(256) k_ffs(const FFArgs A) {
  extern __shared__ unsigned char smem[];
  const int SB = A.sb, AL = A.act_len, D = A.dmax, nl = A.nl, n = A.n, npar = A.npar;
  R* act = reinterpret_cast<R*>(smem);  // SB x AL
  R* db0 = act + (size_t)SB * AL;       // SB x D
  R* db1 = db0 + (size_t)SB * D;        // SB x D
  R* rr = db1 + (size_t)SB * D;         // SB
  const R* P = static_cast<const R*>(A.params);
  const int tid = threadIdx.x;
  const int groups = (A.ns + SB - 1) / SB;
  R* part = MODE == kVjp ? static_cast<R*>(A.partial) + (size_t)blockIdx.x * npar : nullptr;
  bool first = true;
  for (int g = blockIdx.x; g < groups; g += gridDim.x) {
    const int s0 = g * SB;
    const int cnt = min(SB, A.ns - s0);
    const R* act_in = static_cast<const R*>(A.act_in);
    if (MODE == kVjp && act_in) {
      for (int q = tid; q < cnt * AL; q += 256)
        act[q] = act_in[(size_t)s0 * AL + q];
    } else if (A.flip_B > 0) {  // neighbour configurations, as the C4 kernel
      for (int q = tid; q < cnt * n; q += 256) {
        const int s = q / n, i = q % n, cfg = s0 + s;
        int k, l;
        bond_sites(A.flip_kind, n, cfg % A.flip_B, k, l);
        const bool up = A.sigma[(size_t)(cfg / A.flip_B) * n + i] > 0;
        act[(size_t)s * AL + i] = (up != (i == k || i == l)) ? (R)1 : (R)-1;
      }
    } else {
      for (int q = tid; q < cnt * n; q += 256) {
        const int s = q / n, i = q % n;
        act[(size_t)s * AL + i] = A.sigma[(size_t)(s0 + s) * n + i] > 0 ? (R)1 : (R)-1;
      }
    }
    if (MODE == kVjp)
      for (int q = tid; q < cnt; q += 256) rr[q] = static_cast<const R*>(A.r)[s0 + q];
    __syncthreads();
    if (!(MODE == kVjp && act_in)) {
      for (int li = 0; li < nl; ++li) {
        const FFLayerDev L = A.layers[li];
        layer_forward<R>(L, P, act, AL, cnt);
        __syncthreads();
      }
    }
    if (MODE != kVjp && tid < cnt)
      static_cast<R*>(A.out)[s0 + tid] = act[(size_t)tid * AL + AL - 1];
    if (MODE == kFwd) {
      R* ao = static_cast<R*>(A.act_out);
      if (ao)
        for (int q = tid; q < cnt * AL; q += 256) ao[(size_t)s0 * AL + q] = act[q];
      __syncthreads();
      continue;
    }
    // FeedForward::backward (:186-205): dout = 1 at the output
    R* dcur = db0;
    R* dnext = db1;
    if (tid < cnt) dcur[(size_t)tid * D] = (R)1;
    __syncthreads();
    R* const jac = static_cast<R*>(A.jac);
    for (int li = nl - 1; li >= 0; --li) {
      const FFLayerDev L = A.layers[li];
      if (L.npar > 0) {
        if (MODE == kJac) {
          for (int q = tid; q < cnt * L.npar; q += 256) {
            const int s = q / L.npar, c = q % L.npar;
            jac[(size_t)(s0 + s) * npar + L.p_off + c] =
                layer_der<R>(L, act + (size_t)s * AL + L.in_off, dcur + (size_t)s * D, c);
          }
        } else {
          for (int c = tid; c < L.npar; c += 256) {
            R v = (R)0;
            for (int s = 0; s < cnt; ++s)
              v += rr[s] * layer_der<R>(L, act + (size_t)s * AL + L.in_off, dcur + (size_t)s * D, c);
            R* dst = part + L.p_off + c;
            *dst = first ? v : *dst + v;
          }
        }
      }
      if (li > 0) layer_din<R>(L, P, act, AL, dcur, dnext, D, cnt);
      __syncthreads();
      R* t = dcur;
      dcur = dnext;
      dnext = t;
    }
    first = false;
  }
}

template <typename R>
__global__ void k_ffs_reduce(const R* __restrict__ part, int blocks, int npar, R* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= npar) return;
  R v = (R)0;
  for (int b = 0; b < blocks; ++b) v += part[(size_t)b * npar + c];
  out[c] = v;
}

template <typename R>
int launch_r(const FFArgs& a, int mode, hipStream_t st) {
  const size_t lds = ffs_lds_bytes(a.act_len, a.dmax, a.sb, sizeof(R) == 8);
  const int groups = (a.ns + a.sb - 1) / a.sb;
  if (mode == kFwd) {
    ProfGuard g("ff_forward", st);
    hipLaunchKernelGGL((k_ffs<R, kFwd>), dim3(groups), dim3(256), lds, st, a);
  } else if (mode == kJac) {
    ProfGuard g("ff_jacobian", st);
    hipLaunchKernelGGL((k_ffs<R, kJac>), dim3(groups), dim3(256), lds, st, a);
  } else {
    // every block owns at least one group, so each writes its partial row
    const int cap = a.vjp_blocks > 0 ? a.vjp_blocks : 512;
    const int blocks = groups < cap ? groups : cap;
    ProfGuard g("ff_vjp", st);
    hipLaunchKernelGGL((k_ffs<R, kVjp>), dim3(blocks), dim3(256), lds, st, a);
    hipLaunchKernelGGL(k_ffs_reduce<R>, dim3((a.npar + 255) / 256), dim3(256), 0, st,
                       static_cast<const R*>(a.partial), blocks, a.npar, static_cast<R*>(a.grad));
  }
  return YQ_OK;
}

}  // namespace

size_t ffs_lds_bytes(int act_len, int dmax, int sb, int use_double) {
  const size_t rs = use_double ? sizeof(double) : sizeof(float);
  return rs * (size_t)sb * ((size_t)act_len + 2 * (size_t)dmax + 1);
}

int launch_ffs(const FFArgs& a, int mode, int use_double, hipStream_t st) {
  return use_double ? launch_r<double>(a, mode, st) : launch_r<float>(a, mode, st);
}

}  // namespace yq
