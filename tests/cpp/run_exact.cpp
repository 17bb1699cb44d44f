// Exact-sum SR through the concept layer, written the way
// Examples/RBMTFIExact.cpp drives the reference: RunRBMExact<complex<double>>
// on the TFI chain over BasisFull.  Exits non-zero unless the final energy is
// within 0.2% of -2/sin(pi/2N), and checks ExactSampler's draws against the
// basis weights (mean of sigma_0 sigma_1 within 5 sigma).
#include <cmath>
#include <complex>
#include <cstdio>
#include <sstream>

#include "yannq_amd/yannq_amd.hpp"

using namespace yannq_amd;

int main() {
  const int N = 8, alpha = 1;
  Device dev(0);
  std::ostringstream log;
  RunRBMExact<std::complex<double>> runner(dev, N, alpha, true, log);
  runner.setSeed(7);
  runner.initializeRandom(0.01);
  runner.setIterParams(200, 0);
  runner.setOptimizer("SGD", 0.05, 0.0);
  TFIsing ham(N, -1.0, -1.0);
  double last = 0.0;
  auto callback = [&](int ll, double E, double nv) {
    last = E;
    if (ll % 50 == 0) std::printf("%d\t%.10f\t%.4e\n", ll, E, nv);
  };
  runner.run(callback, BasisFull{N}, ham);
  const double e0 = -2.0 / std::sin(M_PI / (2 * N));
  std::printf("final %.10f exact %.10f\n", last, e0);
  if (std::fabs(last - e0) / std::fabs(e0) > 2e-3) return 1;

  // ExactSampler: <sigma_0 sigma_1> from draws vs the exact weighted sum
  auto& qs = runner.getQs();
  SRMatExact<RBM<std::complex<double>>, TFIsing> srex(qs, BasisFull{N}, ham);
  srex.constructExact();
  std::vector<double> p(1u << N);
  detail::check(qs.ctx(), yq_exact_probs(qs.ctx(), p.data()));
  double zz = 0.0;
  for (uint32_t v = 0; v < p.size(); ++v) zz += p[v] * ((v & 1) == ((v >> 1) & 1) ? 1.0 : -1.0);
  ExactSampler<RBM<std::complex<double>>> sampler(qs, BasisFull{N});
  sampler.setSeed(11);
  const int draws = 100000;
  const auto idx = sampler.sample(draws);
  double m = 0.0;
  for (int i : idx) m += ((i & 1) == ((i >> 1) & 1)) ? 1.0 : -1.0;
  m /= draws;
  const double sd = std::sqrt((1.0 - zz * zz) / draws);
  std::printf("zz exact %.6f sampled %.6f (sd %.2e)\n", zz, m, sd);
  return std::fabs(m - zz) < 5 * sd ? 0 : 1;
}
