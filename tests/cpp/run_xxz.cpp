// A complex RBM through the header-only concept layer, written the way
// Examples/RBMXXZ.cpp drives the reference: RunRBM<std::complex<double>> +
// SamplerMT + SwapSweeper + XXZ(N, 1, Delta, signRule = true), lambda
// schedule setLambda(1.0, 0.9, 1e-4), parallel tempering over 16
// temperatures (createSampler(sweeper, 16, .) as at Examples/RBMXXZ.cpp:81,
// here 8 chain groups) started in the Sz = 0 sector.
// N = 10, Delta = 1: prints "final <E> <E_exact>" and exits non-zero when the
// last 20 epochs average more than 1% above the exact ground-state energy.
#include <cmath>
#include <complex>
#include <cstdio>
#include <sstream>

#include "yannq_amd/yannq_amd.hpp"

using namespace yannq_amd;

int main() {
  const int N = 10, alpha = 2;
  const double e0 = -18.06178541796817;  // exact diagonalisation of this chain
  Device dev(0);
  std::ostringstream log;
  RunRBM<std::complex<double>> runner(dev, N, alpha, true, log);
  runner.setSeed(2024);
  runner.initializeRandom(0.01);
  runner.setLambda(1.0, 0.9, 1e-4);
  runner.setIterParams(199, 0);
  runner.setOptimizer("SGD", 0.02, 0.0);
  runner.setSolverParams(false, 1e-3);
  XXZ ham(N, 1.0, 1.0, true);
  SwapSweeper sweeper(N);
  auto sampler = runner.createSampler(sweeper, 16, 8, 77);
  auto randomizer = [N](auto& re) { return randomSigma(N, N / 2, re); };
  double tail = 0.0;
  int cnt = 0;
  auto callback = [&](int ll, double E, double nv, double cgErr, long smp, long slv) {
    if (ll >= 180) {
      tail += E;
      ++cnt;
    }
    if (ll % 40 == 0) std::printf("%d\t%.8f\t%.4e\t%.2e\t%ld\t%ld\n", ll, E, nv, cgErr, smp, slv);
  };
  runner.run(sampler, callback, randomizer, ham, 256, 20);
  tail /= cnt;
  std::printf("final %.8f %.8f\n", tail, e0);
  // the machine is complex: its parameters come back as std::complex<double>
  const auto p = runner.getQs().getParams();
  double im = 0.0;
  for (const auto& z : p) im += std::abs(z.imag());
  std::printf("sum |Im p| %.4e\n", im);
  return std::fabs(tail - e0) / std::fabs(e0) < 0.01 ? 0 : 1;
}
