// CPU oracle of the general FeedForward stack -- test infrastructure only
// (see oracle.h): the layer loops of the reference in its own summation
// order, in double.  Reference citations are paths under the yannq tree.
#include <math.h>

#include <vector>

#include "oracle.h"

namespace {

double sgn(int8_t s) { return s > 0 ? 1.0 : -1.0; }

// layer kinds and activations: the numbering of or_ff_layer (oracle.h)
enum { kConv = 1, kFC = 2, kAct = 3, kMean = 4 };
enum {
  aIdentity = 0, aLnCosh, aTanh, aWeakTanh, aSigmoid, aReLU, aLeakyReLU, aHardTanh,
  aSoftShrink, aLeakyHardTanh, aSoftSign, aCos
};

// outputDim of a layer for an input of width w (Conv1D.hpp:103-105,
// FullyConnected outputDim, ActivationLayer.hpp:98-101, Mean.hpp:71-75)
long long out_dim(const or_ff_layer& L, long long w) {
  switch (L.kind) {
    case kConv: return (w / L.d / L.a) * L.b;
    case kFC: return L.b;
    case kAct: return w;
    default: return L.c == 0 ? L.b : L.a;
  }
}
long long npar_of(const or_ff_layer& L) {
  if (L.kind == kConv) return (long long)L.a * L.b * L.c + (L.use_bias ? L.b : 0);
  if (L.kind == kFC) return (long long)L.a * L.b + (L.use_bias ? L.b : 0);
  return 0;
}

// activations.hpp: operator() and ApplyJacobian, element by element
double act_value(int a, double prm, double z) {
  switch (a) {
    case aLnCosh: return or_logcosh(z);                     // :83-87
    case aTanh: return prm * tanh(z / prm);                 // :126-128
    case aWeakTanh: return z - tanh(z) / 2.0;               // :168-171
    case aSigmoid: return exp(z) + 1.;                      // :214-219 (inverse() discarded)
    case aReLU: return z > 0.0 ? z : 0.0;                   // :303-307
    case aLeakyReLU: return z > 0 ? z : prm * z;            // :404-408
    case aHardTanh: return fabs(z) > 1.0 ? copysign(1.0, z) : z;  // :451-458
    case aSoftShrink: return fabs(z) < 1.0 ? 0.0 : (z - copysign(1.0, z));  // :500-507
    case aLeakyHardTanh: {                                  // :553-562
      double v = fabs(z) > 1.0 ? copysign(1.0, z) : z;
      v += fabs(z) < 1.0 ? 0.0 : prm * (z - copysign(1.0, z));
      return v;
    }
    case aSoftSign: return z / (fabs(z) + 1.0);             // :605-607
    case aCos: return prm * cos(z * M_PI / prm);            // :651-653
    default: return z;                                      // Identity :47-50
  }
}
double act_jac(int a, double prm, double z, double av, double F) {
  switch (a) {
    case aLnCosh: return F * tanh(z);                          // :93-97
    case aTanh: return F * (1. - av * av / (prm * prm));       // :137-141
    case aWeakTanh: {                                          // :178-184
      const double t = tanh(z);
      return F * (1.0 + t * t) / 2.0;
    }
    case aSigmoid: return (1. - av) * av;                      // :226-230 (F unused)
    case aReLU: return z > 0 ? F : 0.0;                        // :314-320
    case aLeakyReLU: return z > 0 ? F : prm * F;               // :415-421
    case aHardTanh: return fabs(z) > 1.0 ? 0.0 : F;            // :465-471
    case aSoftShrink: return fabs(z) < 1.0 ? 0.0 : F;          // :514-520
    case aLeakyHardTanh: return fabs(z) > 1.0 ? prm * F : F;   // :569-575
    case aSoftSign: {                                          // :614-617
      const double d = 1.0 + fabs(z);
      return (1.0 / (d * d)) * F;
    }
    case aCos: return -M_PI * sin(M_PI * z / prm) * F;         // :660-663
    default: return F;
  }
}

// FeedForward::makeData (FeedForward.hpp:170-184): the input and every
// layer's output
std::vector<std::vector<double>> make_data(int n, int nl, const or_ff_layer* ly, const double* p,
                                           const int8_t* sigma) {
  std::vector<std::vector<double>> data;
  std::vector<double> x((size_t)n);
  for (int i = 0; i < n; ++i) x[(size_t)i] = sgn(sigma[i]);
  data.push_back(x);
  long long poff = 0;
  for (int li = 0; li < nl; ++li) {
    const or_ff_layer& L = ly[li];
    const std::vector<double>& in = data.back();
    const long long w = (long long)in.size();
    std::vector<double> out((size_t)out_dim(L, w), 0.0);
    const double* par = p + poff;
    if (L.kind == kConv) {  // Conv1D::forward, Conv1D.hpp:144-171
      const long long ic_n = L.a, oc_n = L.b, K = L.c, stride = L.d;
      const long long inSize = w / ic_n, outSize = inSize / stride;
      for (long long oc = 0; oc < oc_n; oc++)
        for (long long r = 0; r < outSize; r++)
          for (long long ic = 0; ic < ic_n; ic++)
            for (long long ki = 0; ki < K; ki++)
              out[(size_t)(r + oc * outSize)] +=
                  par[(ki + ic * K) + oc * ic_n * K] *
                  in[(size_t)(((r * stride + ki - K / 2 + inSize) % inSize) + ic * inSize)];
      if (L.use_bias)
        for (long long oc = 0; oc < oc_n; ++oc)
          for (long long i = 0; i < outSize; ++i)
            out[(size_t)(i + oc * outSize)] += par[ic_n * K * oc_n + oc];
    } else if (L.kind == kFC) {  // FullyConnected::forward, :150-157
      for (long long o = 0; o < L.b; ++o) {
        double t = 0.0;
        for (long long i = 0; i < L.a; ++i) t += par[i + o * L.a] * in[(size_t)i];
        out[(size_t)o] = L.use_bias ? par[(long long)L.a * L.b + o] + t : t;
      }
    } else if (L.kind == kAct) {
      for (long long i = 0; i < w; ++i) out[(size_t)i] = act_value(L.a, L.param, in[(size_t)i]);
    } else {  // Mean::forward, Mean.hpp:92-103 (input rows x cols column-major)
      if (L.c == 0) {
        for (long long j = 0; j < L.b; ++j) {
          double s = 0.0;
          for (long long i = 0; i < L.a; ++i) s += in[(size_t)(i + j * L.a)];
          out[(size_t)j] = s / L.a;
        }
      } else {
        for (long long i = 0; i < L.a; ++i) {
          double s = 0.0;
          for (long long j = 0; j < L.b; ++j) s += in[(size_t)(i + j * L.a)];
          out[(size_t)i] = s / L.b;
        }
      }
    }
    poff += npar_of(L);
    data.push_back(out);
  }
  return data;
}

}  // namespace

int or_ffs_check(int n, int nl, const or_ff_layer* ly) {
  long long w = n;
  for (int i = 0; i < nl; ++i) {
    const or_ff_layer& L = ly[i];
    if (L.kind == kConv) {
      if (L.a < 1 || L.b < 1 || L.c < 1 || L.d < 1 || w % L.a) return 1;
      if (L.c / 2 > w / L.a || (w / L.a) / L.d < 1) return 1;
    } else if (L.kind == kFC) {
      if (L.a != w || L.b < 1) return 1;
    } else if (L.kind == kAct) {
      if (L.a < aIdentity || L.a > aCos) return 1;
    } else if (L.kind == kMean) {
      if ((long long)L.a * L.b != w || (L.c != 0 && L.c != 1)) return 1;
    } else {
      return 1;
    }
    w = out_dim(L, w);
  }
  return w == 1 ? 0 : 1;
}

int or_ffs_npar(int nl, const or_ff_layer* ly) {
  long long np = 0;
  for (int i = 0; i < nl; ++i) np += npar_of(ly[i]);
  return (int)np;
}

double or_ffs_forward(int n, int nl, const or_ff_layer* ly, const double* p,
                      const int8_t* sigma) {
  return make_data(n, nl, ly, p, sigma).back()[0];
}

// FeedForward::backward (FeedForward.hpp:186-205): dout = 1, each layer's
// backprop in reverse order writing its parameter derivatives
double or_ffs_jacobian(int n, int nl, const or_ff_layer* ly, const double* p,
                       const int8_t* sigma, double* der) {
  const std::vector<std::vector<double>> data = make_data(n, nl, ly, p, sigma);
  std::vector<long long> poff((size_t)nl + 1, 0);
  for (int i = 0; i < nl; ++i) poff[(size_t)i + 1] = poff[(size_t)i] + npar_of(ly[i]);
  std::vector<double> dout(1, 1.0);
  for (int li = nl - 1; li >= 0; --li) {
    const or_ff_layer& L = ly[li];
    const std::vector<double>& in = data[(size_t)li];
    const std::vector<double>& outv = data[(size_t)li + 1];
    const long long w = (long long)in.size();
    std::vector<double> din((size_t)w, 0.0);
    const double* par = p + poff[(size_t)li];
    double* d = der + poff[(size_t)li];
    if (L.kind == kConv) {  // Conv1D::backprop, Conv1D.hpp:173-225
      const long long ic_n = L.a, oc_n = L.b, K = L.c, stride = L.d;
      const long long inSize = w / ic_n, outSize = inSize / stride;
      for (long long oc = 0; oc < oc_n; oc++)
        for (long long r = 0; r < outSize; r++)
          for (long long ic = 0; ic < ic_n; ic++)
            for (long long ki = 0; ki < K; ki++)
              din[(size_t)(((r * stride + ki - K / 2 + inSize) % inSize) + ic * inSize)] +=
                  par[(ki + ic * K) + oc * ic_n * K] * dout[(size_t)(r + oc * outSize)];
      const long long nk = ic_n * K * oc_n;
      for (long long q = 0; q < nk; ++q) d[q] = 0.0;
      for (long long oc = 0; oc < oc_n; oc++)
        for (long long r = 0; r < outSize; r++)
          for (long long ic = 0; ic < ic_n; ic++)
            for (long long ki = 0; ki < K; ki++)
              d[(ki + ic * K) + oc * ic_n * K] +=
                  in[(size_t)(((r * stride + ki - K / 2 + inSize) % inSize) + ic * inSize)] *
                  dout[(size_t)(r + oc * outSize)];
      if (L.use_bias) {
        for (long long oc = 0; oc < oc_n; oc++) d[nk + oc] = 0.0;
        for (long long oc = 0; oc < oc_n; oc++)
          for (long long r = 0; r < outSize; r++) d[nk + oc] += dout[(size_t)(r + oc * outSize)];
      }
    } else if (L.kind == kFC) {  // FullyConnected::backprop, :160-179
      for (long long o = 0; o < L.b; ++o)
        for (long long i = 0; i < L.a; ++i) d[i + o * L.a] = in[(size_t)i] * dout[(size_t)o];
      if (L.use_bias)
        for (long long o = 0; o < L.b; ++o) d[(long long)L.a * L.b + o] = dout[(size_t)o];
      for (long long i = 0; i < L.a; ++i) {
        double t = 0.0;
        for (long long o = 0; o < L.b; ++o) t += par[i + o * L.a] * dout[(size_t)o];
        din[(size_t)i] = t;
      }
    } else if (L.kind == kAct) {  // ActivationLayer::backprop, ActivationLayer.hpp:108-114
      for (long long i = 0; i < w; ++i)
        din[(size_t)i] = act_jac(L.a, L.param, in[(size_t)i], outv[(size_t)i], dout[(size_t)i]);
    } else {  // Mean::backprop, Mean.hpp:106-125
      for (long long j = 0; j < L.b; ++j)
        for (long long i = 0; i < L.a; ++i)
          din[(size_t)(i + j * L.a)] =
              L.c == 0 ? dout[(size_t)j] / L.a : dout[(size_t)i] / L.b;
    }
    dout = din;
  }
  return data.back()[0];
}
