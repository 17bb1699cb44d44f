// CPU oracle, AmplitudePhase machine -- test infrastructure (see oracle.h).
// psi(sigma) = sqrt(RBM(sigma)) exp(i pi FF(sigma)) with a real RBM as the
// amplitude and the FeedForward network as the phase, restated in double
// precision with libm in the reference's order:
//   Machines/AmplitudePhase.hpp (coeff, logDeriv, parameter concatenation),
//   States/AmplitudePhaseState.hpp (logRatio), GroundState/NGDExact.hpp
//   (exact energy, gradient and the amplitude / phase metric blocks).
#include <math.h>

#include <algorithm>
#include <complex>
#include <vector>

#include "functors.h"
#include "oracle.h"

namespace {
using cd = std::complex<double>;
}  // namespace

extern "C" {

// AmplitudePhase.hpp:127-138: log coeff = log RBM coeff / 2 + i pi FF(sigma)
void or_ap_log_psi(int n, int m, const double* W, const double* a, const double* b,
                   const or_ff_desc* d, const double* ffp, const int8_t* sigma, double* out2) {
  std::vector<double> th(m);
  or_theta(n, m, W, b, sigma, th.data());
  out2[0] = or_log_coeff(n, m, a, sigma, th.data()) / 2.0;
  out2[1] = M_PI * or_ff_forward(d, ffp, sigma);
}

// AmplitudePhaseState.hpp:57-120: Re = amplitude logRatio / 2,
// Im = pi (FF(flipped) - FF(sigma)); l < 0 flips k only
void or_ap_log_ratio(int n, int m, const double* W, const double* a, const or_ff_desc* d,
                     const double* ffp, const int8_t* sigma, const double* theta, int k, int l,
                     double* out2) {
  const double lr = l < 0 ? or_log_ratio1(n, m, W, a, sigma, theta, k)
                          : or_log_ratio2(n, m, W, a, sigma, theta, k, l);
  std::vector<int8_t> flipped(sigma, sigma + n);
  flipped[k] = (int8_t)-flipped[k];
  if (l >= 0) flipped[l] = (int8_t)-flipped[l];
  out2[0] = lr / 2.0;
  out2[1] = M_PI * (or_ff_forward(d, ffp, flipped.data()) - or_ff_forward(d, ffp, sigma));
}

// The Hamiltonians' functors with ratio = exp(logRatio) of the state above
void or_ap_eloc(const or_ham* H, int n, int m, const double* W, const double* a,
                const or_ff_desc* d, const double* ffp, const int8_t* s, const double* theta,
                double* out2) {
  auto ratio1 = [&](int k) {
    double r[2];
    or_ap_log_ratio(n, m, W, a, d, ffp, s, theta, k, -1, r);
    return std::exp(cd(r[0], r[1]));
  };
  auto ratio2 = [&](int k, int l) {
    double r[2];
    or_ap_log_ratio(n, m, W, a, d, ffp, s, theta, k, l, r);
    return std::exp(cd(r[0], r[1]));
  };
  const cd e = oracle_detail::eloc_functor(H, s, ratio1, ratio2);
  out2[0] = e.real();
  out2[1] = e.imag();
}

// AmplitudePhase.hpp:42-58: the real rows [logDerivAmp | logDerivPhase] =
// [RBM logDeriv / 2 | pi FF backward]; logDeriv is [first | i * second]
void or_ap_logderiv(int n, int m, int use_bias, const or_ff_desc* d, const double* ffp,
                    const int8_t* sigma, const double* theta, double* out) {
  const int dimA = n * m + (use_bias ? n + m : 0);
  or_logderiv(n, m, use_bias, sigma, theta, out);
  for (int q = 0; q < dimA; ++q) out[q] = out[q] / 2.0;
  const int dimP = or_ff_npar(d);
  or_ff_jacobian(d, ffp, sigma, out + dimA);
  for (int q = 0; q < dimP; ++q) out[dimA + q] = M_PI * out[dimA + q];
}

// NGDExact.hpp:137-184 over a sorted basis of B configurations (bit i set
// <=> sigma_i = -1).  The Hamiltonian is its column functor restricted to
// the basis (ED/ConstructSparseMat.hpp:32-57; entries outside the basis are
// dropped).  Outputs: E, olocAmp (dimA), olocPhase (dimP), grad (dimA +
// dimP), corrMatAmp (dimA^2) and corrMatPhase (dimP^2), column-major.
// Returns 0, or -1 when a matrix element leaves the basis.
int or_ap_ngd_exact(const or_ham* H, int n, int m, int use_bias, const double* rbm_params,
                    const or_ff_desc* d, const double* ffp, const uint32_t* basis, int B,
                    double* E, double* olocAmp, double* olocPhase, double* grad, double* Samp,
                    double* Sphase) {
  const int dimA = n * m + (use_bias ? n + m : 0), dimP = or_ff_npar(d);
  const double* W = rbm_params;
  std::vector<double> zeros_a(n, 0.0), zeros_b(m, 0.0);
  const double* a = use_bias ? rbm_params + (size_t)n * m : zeros_a.data();
  const double* b = use_bias ? rbm_params + (size_t)n * m + n : zeros_b.data();
  // getPsi(qs, basis, true): coeff = sqrt(RBM coeff) exp(i pi FF), normalised
  std::vector<cd> st(B);
  std::vector<std::vector<int8_t>> sig(B, std::vector<int8_t>(n));
  std::vector<std::vector<double>> th(B, std::vector<double>(m));
  for (int s = 0; s < B; ++s) {
    for (int i = 0; i < n; ++i) sig[s][i] = ((basis[s] >> i) & 1u) ? -1 : 1;
    or_theta(n, m, W, b, sig[s].data(), th[s].data());
    const double amp = exp(or_log_coeff(n, m, a, sig[s].data(), th[s].data()));
    st[s] = std::sqrt(amp) * std::exp(cd(0.0, M_PI * or_ff_forward(d, ffp, sig[s].data())));
  }
  double nrm = 0.0;
  for (const cd& v : st) nrm += std::norm(v);
  nrm = sqrt(nrm);
  for (cd& v : st) v /= nrm;
  // k = H st
  std::vector<cd> k(B, 0.0);
  std::vector<uint32_t> rows(16 * n + 16);
  std::vector<double> vals(16 * n + 16);
  for (int c = 0; c < B; ++c) {
    const int cnt = or_ham_column(H, basis[c], rows.data(), vals.data(), (int)vals.size());
    for (int q = 0; q < cnt; ++q) {
      const uint32_t* it = std::lower_bound(basis, basis + B, rows[q]);
      if (it == basis + B || *it != rows[q]) {
        if (vals[q] != 0.0) return -1;
        continue;
      }
      k[it - basis] += vals[q] * st[c];
    }
  }
  cd t = 0.0;
  for (int s = 0; s < B; ++s) t += std::conj(st[s]) * k[s];
  *E = t.real();
  // deltas (rows), |st|^2 weighted means
  std::vector<double> rowsA((size_t)B * dimA), rowsP((size_t)B * dimP), row(dimA + dimP);
  std::vector<double> p(B);
  for (int s = 0; s < B; ++s) {
    or_ap_logderiv(n, m, use_bias, d, ffp, sig[s].data(), th[s].data(), row.data());
    std::copy(row.begin(), row.begin() + dimA, rowsA.begin() + (size_t)s * dimA);
    std::copy(row.begin() + dimA, row.end(), rowsP.begin() + (size_t)s * dimP);
    p[s] = std::norm(st[s]);
  }
  for (int c = 0; c < dimA; ++c) olocAmp[c] = 0.0;
  for (int c = 0; c < dimP; ++c) olocPhase[c] = 0.0;
  for (int s = 0; s < B; ++s) {
    for (int c = 0; c < dimA; ++c) olocAmp[c] += p[s] * rowsA[(size_t)s * dimA + c];
    for (int c = 0; c < dimP; ++c) olocPhase[c] += p[s] * rowsP[(size_t)s * dimP + c];
  }
  // k <- conj(st) k; grad = [2 deltasAmp^T Re k - 2 E olocAmp | 2 deltasPhase^T Im k]
  for (int s = 0; s < B; ++s) k[s] = std::conj(st[s]) * k[s];
  for (int c = 0; c < dimA; ++c) {
    double acc = 0.0;
    for (int s = 0; s < B; ++s) acc += rowsA[(size_t)s * dimA + c] * k[s].real();
    grad[c] = 2.0 * acc - 2.0 * *E * olocAmp[c];
  }
  for (int c = 0; c < dimP; ++c) {
    double acc = 0.0;
    for (int s = 0; s < B; ++s) acc += rowsP[(size_t)s * dimP + c] * k[s].imag();
    grad[dimA + c] = 2.0 * acc;
  }
  // corrMat blocks: deltas^T diag(p) deltas - oloc oloc^T
  auto corr = [&](const std::vector<double>& R, int dim, const double* ol, double* S) {
    for (int j = 0; j < dim; ++j)
      for (int i = 0; i < dim; ++i) {
        double acc = 0.0;
        for (int s = 0; s < B; ++s)
          acc += R[(size_t)s * dim + i] * p[s] * R[(size_t)s * dim + j];
        S[(size_t)j * dim + i] = acc - ol[i] * ol[j];
      }
  };
  if (Samp) corr(rowsA, dimA, olocAmp, Samp);
  if (Sphase) corr(rowsP, dimP, olocPhase, Sphase);
  return 0;
}

}  // extern "C"
