// CPU oracle, complex RBM part -- test infrastructure (see oracle.h).
// Restatement with std::complex<double> in the reference's order, and the
// replay of the complex sampler kernel.
#include <math.h>
#include <string.h>

#include <complex>
#include <vector>

#include "../yannq_amd/csrc/yq_math.h"
#include "functors.h"
#include "oracle.h"

namespace {
using cd = std::complex<double>;
inline double sgn(int8_t s) { return s > 0 ? 1.0 : -1.0; }
inline cd at(const double* p, size_t i) { return cd(p[2 * i], p[2 * i + 1]); }
inline void put(double* p, size_t i, cd v) {
  p[2 * i] = v.real();
  p[2 * i + 1] = v.imag();
}
// Utilities/Utility.hpp:28-37
cd logcosh_c(cd x) {
  const double xr = x.real(), xi = x.imag();
  cd res = or_logcosh(xr);
  res += std::log(cd(std::cos(xi), std::tanh(xr) * std::sin(xi)));
  return res;
}
}  // namespace

extern "C" {

void or_c_logcosh(double re, double im, double* out2) { put(out2, 0, logcosh_c(cd(re, im))); }

// Machines/RBM.hpp:137-141
void or_c_theta(int n, int m, const double* W, const double* b, const int8_t* sigma,
                double* theta) {
  for (int j = 0; j < m; ++j) {
    cd acc = 0.0;
    for (int i = 0; i < n; ++i) acc += at(W, (size_t)i * m + j) * sgn(sigma[i]);
    put(theta, j, acc + at(b, j));
  }
}

// States/RBMState.hpp:48-63
void or_c_log_ratio1(int n, int m, const double* W, const double* a, const int8_t* sigma,
                     const double* theta, int k, double* out2) {
  (void)n;
  cd res = -2.0 * at(a, k) * sgn(sigma[k]);
  for (int j = 0; j < m; ++j)
    res += logcosh_c(at(theta, j) - 2.0 * sgn(sigma[k]) * at(W, (size_t)k * m + j)) -
           logcosh_c(at(theta, j));
  put(out2, 0, res);
}

// States/RBMState.hpp:78-93
void or_c_log_ratio2(int n, int m, const double* W, const double* a, const int8_t* sigma,
                     const double* theta, int k, int l, double* out2) {
  (void)n;
  cd res = -2.0 * at(a, k) * sgn(sigma[k]) - 2.0 * at(a, l) * sgn(sigma[l]);
  for (int j = 0; j < m; ++j) {
    const cd t = at(theta, j) - 2.0 * sgn(sigma[k]) * at(W, (size_t)k * m + j) -
                 2.0 * sgn(sigma[l]) * at(W, (size_t)l * m + j);
    res += logcosh_c(t) - logcosh_c(at(theta, j));
  }
  put(out2, 0, res);
}

// States/RBMState.hpp:108-126 (logRatio(vector)) for complex parameters
void or_c_log_ratio_vec(int n, int m, const double* W, const double* a, const int8_t* sigma,
                        const double* theta, int k, const int* v, double* out2) {
  (void)n;
  cd res = 0.0;
  for (int e = 0; e < k; ++e) res -= 2.0 * at(a, v[e]) * sgn(sigma[v[e]]);
  for (int j = 0; j < m; ++j) {
    cd t = at(theta, j);
    for (int e = 0; e < k; ++e) t -= 2.0 * sgn(sigma[v[e]]) * at(W, (size_t)v[e] * m + j);
    res += logcosh_c(t) - logcosh_c(at(theta, j));
  }
  put(out2, 0, res);
}

// Machines/RBM.hpp:367-378
void or_c_log_coeff(int n, int m, const double* a, const int8_t* sigma, const double* theta,
                    double* out2) {
  cd s = 0.0;
  for (int i = 0; i < n; ++i) s += at(a, i) * sgn(sigma[i]);
  for (int j = 0; j < m; ++j) s += logcosh_c(at(theta, j));
  put(out2, 0, s);
}

// Machines/RBM.hpp:390-406
void or_c_logderiv(int n, int m, int use_bias, const int8_t* sigma, const double* theta,
                   double* out) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < m; ++j) put(out, (size_t)i * m + j, sgn(sigma[i]) * std::tanh(at(theta, j)));
  if (!use_bias) return;
  for (int i = 0; i < n; ++i) put(out, (size_t)n * m + i, sgn(sigma[i]));
  for (int j = 0; j < m; ++j) put(out, (size_t)n * m + n + j, std::tanh(at(theta, j)));
}

// TFIsing.hpp:28-39, XXZ.hpp:37-49, J1J2.hpp:35-55 with ratio = exp(logRatio)
void or_c_eloc(const or_ham* H, int n, int m, const double* W, const double* a,
               const int8_t* s, const double* theta, double* out2) {
  auto ratio1 = [&](int k) {
    double r[2];
    or_c_log_ratio1(n, m, W, a, s, theta, k, r);
    return std::exp(cd(r[0], r[1]));
  };
  auto ratio2 = [&](int k, int l) {
    double r[2];
    or_c_log_ratio2(n, m, W, a, s, theta, k, l, r);
    return std::exp(cd(r[0], r[1]));
  };
  put(out2, 0, oracle_detail::eloc_functor(H, s, ratio1, ratio2));
}

// GroundState/utils.hpp:13-51 and SRMat.hpp:89-103, 142-149 (adjoint products)
void or_c_sr_build(int ns, int n, int m, int use_bias, const int8_t* sigma, const double* theta,
                   const double* eloc, const double* weights, double* O, double* omean,
                   double* F, double* S) {
  const size_t dim = use_bias ? (size_t)n * m + n + m : (size_t)n * m;
  for (int s = 0; s < ns; ++s)
    or_c_logderiv(n, m, use_bias, sigma + (size_t)s * n, theta + (size_t)s * m * 2,
                  O + (size_t)s * dim * 2);
  auto w = [&](int s) { return weights ? weights[s] : 1.0 / ns; };
  for (size_t c = 0; c < dim; ++c) {
    cd acc = 0.0;
    for (int s = 0; s < ns; ++s) acc += w(s) * at(O, (size_t)s * dim + c);
    put(omean, c, acc);
  }
  for (int s = 0; s < ns; ++s)
    for (size_t c = 0; c < dim; ++c)
      put(O, (size_t)s * dim + c, at(O, (size_t)s * dim + c) - at(omean, c));
  for (size_t c = 0; c < dim; ++c) {
    cd acc = 0.0;
    for (int s = 0; s < ns; ++s) acc += w(s) * std::conj(at(O, (size_t)s * dim + c)) * at(eloc, s);
    put(F, c, acc);
  }
  if (!S) return;
  for (size_t c = 0; c < dim; ++c)
    for (size_t d = 0; d < dim; ++d) {
      cd acc = 0.0;
      for (int s = 0; s < ns; ++s)
        acc += w(s) * std::conj(at(O, (size_t)s * dim + c)) * at(O, (size_t)s * dim + d);
      put(S, c * dim + d, acc);
    }
}

double rp_logcosh_re_d(double x, double y) { return yq::logcosh_re_d(x, y); }
float rp_logcosh_re_f(float x, float y) { return yq::logcosh_re_f(x, y); }
void rp_logcosh_c_d(double x, double y, double* out2) { yq::logcosh_c_d(x, y, out2, out2 + 1); }
void rp_logcosh_c_f(float x, float y, float* out2) { yq::logcosh_c_f(x, y, out2, out2 + 1); }
void rp_tanh_c_d(double x, double y, double* out2) { yq::tanh_c_d(x, y, out2, out2 + 1); }

}  // extern "C"

namespace {
// kernel replay: complex theta (re, im) in working precision R
template <typename R>
R lc_re(R x, R y);
template <>
double lc_re<double>(double x, double y) {
  return yq::logcosh_re_d(x, y);
}
template <>
float lc_re<float>(float x, float y) {
  return yq::logcosh_re_f(x, y);
}

template <typename R>
void replay_chain_c(const rp_cfg* cfg, uint32_t stream, const double* W, const double* a,
                    const double* b, int8_t* sig, uint64_t* ctr, uint64_t* acc, int n_sweeps,
                    int n_discard, int C, int chain, int8_t* out_sigma, double* out_theta) {
  const int n = cfg->n, m = cfg->m, L = cfg->lanes;
  const bool mixed = cfg->precision == 1;
  std::vector<R> Wr((size_t)n * m), Wi((size_t)n * m), ar(n), tr(m), ti(m), lc(m), ur(m),
      ui(m), lt(m), part(L);
  for (size_t q = 0; q < (size_t)n * m; ++q) {
    Wr[q] = (R)W[2 * q];
    Wi[q] = (R)W[2 * q + 1];
  }
  for (int i = 0; i < n; ++i) ar[i] = (R)a[2 * i];
  std::vector<double> t64(2 * (size_t)m);
  auto refresh = [&]() {
    for (int j = 0; j < m; ++j) {
      double sr = 0.0, si = 0.0;
      for (int i = 0; i < n; ++i) {
        const double wr = W[2 * ((size_t)i * m + j)], wi = W[2 * ((size_t)i * m + j) + 1];
        sr = sig[i] > 0 ? sr + wr : sr - wr;
        si = sig[i] > 0 ? si + wi : si - wi;
      }
      t64[2 * j] = sr + b[2 * j];
      t64[2 * j + 1] = si + b[2 * j + 1];
      tr[j] = (R)t64[2 * j];
      ti[j] = (R)t64[2 * j + 1];
      lc[j] = lc_re<R>(tr[j], ti[j]);
    }
  };
  refresh();
  const int nprop = n * cfg->sweep_factor;
  for (int sw = 0; sw < n_discard + n_sweeps; ++sw) {
    if (mixed && sw > 0) refresh();
    for (int p = 0; p < nprop; ++p) {
      const yq::u32x4 x = yq::draw(cfg->seed, (*ctr)++, stream, yq::kDomSweep);
      const int k = (int)yq::index_from(x.x, (uint32_t)n);
      int l = -1;
      R cst;
      if (cfg->kind == 1) {
        l = (int)yq::index_from(x.y, (uint32_t)n);
        if (sig[k] == sig[l]) continue;
        cst = ((R)-2 * ar[k]) * (R)sig[k] - ((R)2 * ar[l]) * (R)sig[l];
      } else {
        cst = ((R)-2 * ar[k]) * (R)sig[k];
      }
      for (int q = 0; q < L; ++q) part[q] = (R)0;
      for (int j = 0; j < m; ++j) {
        const R tk = (R)(2 * sig[k]);
        R vr = tr[j] - tk * Wr[(size_t)k * m + j];
        R vi = ti[j] - tk * Wi[(size_t)k * m + j];
        if (l >= 0) {
          const R tl = (R)(2 * sig[l]);
          vr = vr - tl * Wr[(size_t)l * m + j];
          vi = vi - tl * Wi[(size_t)l * m + j];
        }
        ur[j] = vr;
        ui[j] = vi;
        lt[j] = lc_re<R>(vr, vi);
        part[j % L] = part[j % L] + (lt[j] - lc[j]);
      }
      for (int s = L / 2; s >= 1; s /= 2)
        for (int q = 0; q < s; ++q) part[q] = part[q] + part[q + s];
      const R delta = cst + part[0];
      const bool ok = mixed ? yq::accept_f((float)cfg->beta, (float)delta, yq::u01_float(x.z))
                            : yq::accept_d(cfg->beta, (double)delta, yq::u01_double(x.z, x.w));
      if (ok) {
        for (int j = 0; j < m; ++j) {
          tr[j] = ur[j];
          ti[j] = ui[j];
          lc[j] = lt[j];
        }
        sig[k] = -sig[k];
        if (l >= 0) sig[l] = -sig[l];
        ++(*acc);
      }
    }
    if (sw >= n_discard) {
      const size_t row = (size_t)(sw - n_discard) * C + chain;
      memcpy(out_sigma + row * n, sig, n);
      if (mixed) refresh();
      for (int j = 0; j < m; ++j) {
        out_theta[2 * (row * m + j)] = (double)tr[j];
        out_theta[2 * (row * m + j) + 1] = (double)ti[j];
      }
    }
  }
}
}  // namespace

namespace {
template <typename R>
void lc_c(R x, R y, R* re, R* im);
template <>
void lc_c<double>(double x, double y, double* re, double* im) {
  yq::logcosh_c_d(x, y, re, im);
}
template <>
void lc_c<float>(float x, float y, float* re, float* im) {
  yq::logcosh_c_f(x, y, re, im);
}
inline void sc(double x, double* s, double* c) { yq::sincos_d(x, s, c); }
inline void sc(float x, float* s, float* c) { yq::sincos_f(x, s, c); }

// One sample of the complex E_loc kernels (k_eloc_ratio_c / k_eloc_combine_c):
// theta rounded to R, the per-bond log-ratio summed over j in index order,
// ratio = exp(Re) (cos Im + i sin Im), then the functor's accumulation in
// double.  The phase uses the host atan2, so agreement with the device is to
// a few ulp rather than bit for bit.
template <typename R>
std::complex<double> c_eloc_one(const or_ham* H, int n, int m, const std::vector<R>& Wr,
                                const std::vector<R>& Wi, const std::vector<R>& ar,
                                const std::vector<R>& ai, const int8_t* s, const double* th) {
  std::vector<R> tr(m), ti(m), lr(m), li(m);
  for (int j = 0; j < m; ++j) {
    tr[j] = (R)th[2 * j];
    ti[j] = (R)th[2 * j + 1];
    lc_c<R>(tr[j], ti[j], &lr[j], &li[j]);
  }
  auto ratio = [&](int k, int l) -> std::complex<double> {
    if (l >= 0 && s[k] == s[l]) return 0.0;
    const R twok = s[k] > 0 ? (R)2 : (R)-2;
    R rr = ((R)-2 * ar[k]) * (R)s[k], ri = ((R)-2 * ai[k]) * (R)s[k];
    R twol = (R)0;
    if (l >= 0) {
      twol = s[l] > 0 ? (R)2 : (R)-2;
      rr = rr - ((R)2 * ar[l]) * (R)s[l];
      ri = ri - ((R)2 * ai[l]) * (R)s[l];
    }
    for (int j = 0; j < m; ++j) {
      R x = tr[j] - twok * Wr[(size_t)k * m + j];
      R y = ti[j] - twok * Wi[(size_t)k * m + j];
      if (l >= 0) {
        x = x - twol * Wr[(size_t)l * m + j];
        y = y - twol * Wi[(size_t)l * m + j];
      }
      R a, b;
      lc_c<R>(x, y, &a, &b);
      rr = rr + (a - lr[j]);
      ri = ri + (b - li[j]);
    }
    R sn, cs;
    sc(ri, &sn, &cs);
    const R mag = yq::Math<R>::exp(rr);
    return std::complex<double>((double)(mag * cs), (double)(mag * sn));
  };
  double er = 0.0, ei = 0.0;
  auto add = [&](double c, std::complex<double> r) {
    er += c * r.real();
    ei += c * r.imag();
  };
  if (H->kind == OR_TFI) {
    for (int i = 0; i < n; ++i) {
      er += H->J * (double)s[i] * (double)s[(i + 1) % n];
      add(H->h, ratio(i, -1));
    }
  } else if (H->kind == OR_XXZ) {
    const double sign = H->sign_rule ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) {
      const int zz = (int)s[i] * (int)s[(i + 1) % n];
      er += H->J * H->delta * zz;
      add(sign * H->J * (1 - zz), ratio(i, (i + 1) % n));
    }
  } else {
    const int sign = H->sign_rule ? -1 : 1;
    for (int i = 0; i < n; ++i) {
      const double yy = -(double)((int)s[i] * (int)s[(i + 1) % n]);
      er += -H->J * yy;
      add(H->J * sign * (1.0 + yy), ratio(i, (i + 1) % n));
    }
    for (int i = 0; i < n; ++i) {
      const double yy = -(double)((int)s[i] * (int)s[(i + 2) % n]);
      er += -H->J2 * yy;
      add(H->J2 * (1.0 + yy), ratio(i, (i + 2) % n));
    }
  }
  return std::complex<double>(er, ei);
}

template <typename R>
void c_eloc_all(const or_ham* H, int ns, int n, int m, const double* W, const double* a,
                const int8_t* sigma, const double* theta, double* eloc) {
  std::vector<R> Wr((size_t)n * m), Wi((size_t)n * m), ar(n), ai(n);
  for (size_t q = 0; q < (size_t)n * m; ++q) {
    Wr[q] = (R)W[2 * q];
    Wi[q] = (R)W[2 * q + 1];
  }
  for (int i = 0; i < n; ++i) {
    ar[i] = (R)a[2 * i];
    ai[i] = (R)a[2 * i + 1];
  }
  for (int s = 0; s < ns; ++s) {
    const std::complex<double> e = c_eloc_one<R>(H, n, m, Wr, Wi, ar, ai, sigma + (size_t)s * n,
                                                 theta + 2 * (size_t)s * m);
    eloc[2 * s] = e.real();
    eloc[2 * s + 1] = e.imag();
  }
}
}  // namespace

extern "C" void rp_c_eloc(const or_ham* H, int precision, int ns, int n, int m, const double* W,
                          const double* a, const int8_t* sigma, const double* theta,
                          double* eloc) {
  if (precision == 1)
    c_eloc_all<float>(H, ns, n, m, W, a, sigma, theta, eloc);
  else
    c_eloc_all<double>(H, ns, n, m, W, a, sigma, theta, eloc);
}

namespace {
// Parallel tempering of a complex machine with the kernel's arithmetic
// (SamplerMT.hpp:44-60, 120-153, 180-204; the real machine's replay,
// replay_pt_group in oracle.cpp, with complex theta): slots t = 0 .. T-1 of
// chain group g at beta_t = (T - t) / T, sweeps with Re logcosh, then the
// two swap passes on lp = Re a . sigma + sum_j Re logcosh(theta_j), each
// sum as the lane-strided tree; slot 0 recorded after every sweep past the
// discard phase.
template <typename R>
void replay_pt_group_c(const rp_cfg* cfg, int T, int groups, int g, const uint32_t* stream,
                       const double* W, const double* a, const double* b, int8_t* sigma,
                       uint64_t* ctr, uint64_t* acc, int n_sweeps, int n_discard,
                       int8_t* out_sigma, double* out_theta) {
  const int n = cfg->n, m = cfg->m, L = cfg->lanes;
  const bool mixed = cfg->precision == 1;
  std::vector<R> Wr((size_t)n * m), Wi((size_t)n * m), ar(n), ur(m), ui(m), lt(m), part(L);
  for (size_t q = 0; q < (size_t)n * m; ++q) {
    Wr[q] = (R)W[2 * q];
    Wi[q] = (R)W[2 * q + 1];
  }
  for (int i = 0; i < n; ++i) ar[i] = (R)a[2 * i];
  struct Slot {
    int8_t* sig;
    std::vector<R> tr, ti, lc;
    uint64_t* ctr;
    uint64_t* acc;
    uint32_t stream;
  };
  std::vector<Slot> ch(T);
  auto refresh = [&](Slot& c) {
    for (int j = 0; j < m; ++j) {
      double sr = 0.0, si = 0.0;
      for (int i = 0; i < n; ++i) {
        const double wr = W[2 * ((size_t)i * m + j)], wi = W[2 * ((size_t)i * m + j) + 1];
        sr = c.sig[i] > 0 ? sr + wr : sr - wr;
        si = c.sig[i] > 0 ? si + wi : si - wi;
      }
      c.tr[j] = (R)(sr + b[2 * j]);
      c.ti[j] = (R)(si + b[2 * j + 1]);
      c.lc[j] = lc_re<R>(c.tr[j], c.ti[j]);
    }
  };
  auto tree = [&]() {
    for (int s = L / 2; s >= 1; s /= 2)
      for (int q = 0; q < s; ++q) part[q] = part[q] + part[q + s];
    return part[0];
  };
  for (int q = 0; q < T; ++q) {
    const int c = g * T + q;
    ch[q].sig = sigma + (size_t)c * n;
    ch[q].tr.resize(m);
    ch[q].ti.resize(m);
    ch[q].lc.resize(m);
    ch[q].ctr = ctr + c;
    ch[q].acc = acc + c;
    ch[q].stream = stream[c];
    refresh(ch[q]);
  }
  const int nprop = n * cfg->sweep_factor;
  std::vector<R> lp(T);
  for (int sw = 0; sw < n_discard + n_sweeps; ++sw) {
    for (int q = 0; q < T; ++q) {
      Slot& c = ch[q];
      if (mixed && sw > 0) refresh(c);
      const R beta = (R)((double)(T - q) / T);
      for (int p = 0; p < nprop; ++p) {
        const yq::u32x4 x = yq::draw(cfg->seed, (*c.ctr)++, c.stream, yq::kDomSweep);
        const int k = (int)yq::index_from(x.x, (uint32_t)n);
        int l = -1;
        R cst;
        if (cfg->kind == 1) {
          l = (int)yq::index_from(x.y, (uint32_t)n);
          if (c.sig[k] == c.sig[l]) continue;
          cst = ((R)-2 * ar[k]) * (R)c.sig[k] - ((R)2 * ar[l]) * (R)c.sig[l];
        } else {
          cst = ((R)-2 * ar[k]) * (R)c.sig[k];
        }
        for (int z = 0; z < L; ++z) part[z] = (R)0;
        for (int j = 0; j < m; ++j) {
          const R tk = (R)(2 * c.sig[k]);
          R vr = c.tr[j] - tk * Wr[(size_t)k * m + j];
          R vi = c.ti[j] - tk * Wi[(size_t)k * m + j];
          if (l >= 0) {
            const R tl = (R)(2 * c.sig[l]);
            vr = vr - tl * Wr[(size_t)l * m + j];
            vi = vi - tl * Wi[(size_t)l * m + j];
          }
          ur[j] = vr;
          ui[j] = vi;
          lt[j] = lc_re<R>(vr, vi);
          part[j % L] = part[j % L] + (lt[j] - c.lc[j]);
        }
        const R delta = cst + tree();
        const bool ok = mixed ? yq::accept_f((float)beta, (float)delta, yq::u01_float(x.z))
                              : yq::accept_d((double)beta, (double)delta,
                                             yq::u01_double(x.z, x.w));
        if (ok) {
          for (int j = 0; j < m; ++j) {
            c.tr[j] = ur[j];
            c.ti[j] = ui[j];
            c.lc[j] = lt[j];
          }
          c.sig[k] = -c.sig[k];
          if (l >= 0) c.sig[l] = -c.sig[l];
          ++(*c.acc);
        }
      }
      for (int z = 0; z < L; ++z) part[z] = (R)0;
      for (int i = 0; i < n; ++i) part[i % L] = part[i % L] + ar[i] * (R)c.sig[i];
      const R sa = tree();
      for (int z = 0; z < L; ++z) part[z] = (R)0;
      for (int j = 0; j < m; ++j) part[j % L] = part[j % L] + c.lc[j];
      const R sl = tree();
      lp[q] = sa + sl;
    }
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<int> d(T, 0);
      for (int q = pass; q + 1 < T && (pass == 0 || q + 1 <= T - 2); q += 2) {
        const yq::u32x4 x = yq::draw(cfg->seed, *ch[q].ctr, ch[q].stream, yq::kDomMix);
        const double dbeta = (double)(T - q - 1) / T - (double)(T - q) / T;
        const R dl = lp[q] - lp[q + 1];
        d[q] = mixed ? yq::accept_f((float)dbeta, (float)dl, yq::u01_float(x.z))
                     : yq::accept_d(dbeta, (double)dl, yq::u01_double(x.z, x.w));
      }
      for (int q = 0; q + 1 < T; ++q) {
        if (!d[q]) continue;
        Slot& A = ch[q];
        Slot& B = ch[q + 1];
        for (int i = 0; i < n; ++i) std::swap(A.sig[i], B.sig[i]);
        std::swap(A.tr, B.tr);
        std::swap(A.ti, B.ti);
        for (int j = 0; j < m; ++j) {
          A.lc[j] = lc_re<R>(A.tr[j], A.ti[j]);
          B.lc[j] = lc_re<R>(B.tr[j], B.ti[j]);
        }
        std::swap(lp[q], lp[q + 1]);
      }
    }
    if (sw >= n_discard) {
      const size_t row = (size_t)(sw - n_discard) * groups + g;
      Slot& c = ch[0];
      memcpy(out_sigma + row * n, c.sig, n);
      if (mixed) refresh(c);
      for (int j = 0; j < m; ++j) {
        out_theta[2 * (row * m + j)] = (double)c.tr[j];
        out_theta[2 * (row * m + j) + 1] = (double)c.ti[j];
      }
    }
  }
}
}  // namespace

extern "C" int rp_c_sample_pt(const rp_cfg* cfg, int T, int groups, const uint32_t* stream,
                              const double* W, const double* a, const double* b, int8_t* sigma,
                              uint64_t* ctr, uint64_t* acc, int n_sweeps, int n_discard,
                              int8_t* out_sigma, double* out_theta) {
  if (cfg->lanes <= 0 || (cfg->lanes & (cfg->lanes - 1)) || T < 2 || T % 2) return 1;
  for (int g = 0; g < groups; ++g) {
    if (cfg->precision == 1)
      replay_pt_group_c<float>(cfg, T, groups, g, stream, W, a, b, sigma, ctr, acc, n_sweeps,
                               n_discard, out_sigma, out_theta);
    else
      replay_pt_group_c<double>(cfg, T, groups, g, stream, W, a, b, sigma, ctr, acc, n_sweeps,
                                n_discard, out_sigma, out_theta);
  }
  return 0;
}

extern "C" int rp_c_sample(const rp_cfg* cfg, int C, const uint32_t* stream, const double* W,
                           const double* a, const double* b, int8_t* sigma, uint64_t* ctr,
                           uint64_t* acc, int n_sweeps, int n_discard, int8_t* out_sigma,
                           double* out_theta) {
  if (cfg->lanes <= 0 || (cfg->lanes & (cfg->lanes - 1))) return 1;
  for (int c = 0; c < C; ++c) {
    int8_t* sig = sigma + (size_t)c * cfg->n;
    if (cfg->precision == 1)
      replay_chain_c<float>(cfg, stream[c], W, a, b, sig, ctr + c, acc + c, n_sweeps, n_discard,
                            C, c, out_sigma, out_theta);
    else
      replay_chain_c<double>(cfg, stream[c], W, a, b, sig, ctr + c, acc + c, n_sweeps,
                             n_discard, C, c, out_sigma, out_theta);
  }
  return 0;
}
