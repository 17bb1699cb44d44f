// C1 through the header-only concept layer, written the way
// Examples/RBMTFI.cpp drives the reference: RunRBM + SamplerMT +
// LocalSweeper + TFIsing(N, -1, -h).  Prints "final <E> <E_exact>" and exits
// non-zero when the last 20 epochs average more than 0.5% above the exact
// ground-state energy -2/sin(pi/2N).  Modes: exact (Cholesky), cg (factored
// CG), fisher (Cholesky of the fisher average, setMomentum(0.3, 0.6)).
#include <cmath>
#include <cstdio>
#include <iostream>
#include <sstream>

#include "yannq_amd/yannq_amd.hpp"

using namespace yannq_amd;

int main(int argc, char** argv) {
  const int N = 20, alpha = 1;
  const double h = 1.0;
  const std::string mode = argc > 1 ? argv[1] : "exact";
  const bool useCG = mode == "cg";
  Device dev(0);
  std::ostringstream log;
  RunRBM<double> runner(dev, N, alpha, true, log);
  runner.setSeed(1234);
  runner.initializeRandom(0.01);
  runner.setIterParams(299, 0);
  runner.setOptimizer("SGD", 0.01, 0.0);
  runner.setSolverParams(useCG, 1e-8);
  if (mode == "fisher") runner.setMomentum(0.3, 0.6);
  TFIsing ham(N, -1.0, -h);
  LocalSweeper sweeper(N);
  auto sampler = runner.createSampler(sweeper, 1, 16, 4321);
  auto randomizer = [N](auto& re) { return randomSigma(N, re); };
  double tail = 0.0;
  int cnt = 0;
  auto callback = [&](int ll, double E, double nv, double cgErr, long smp, long slv) {
    if (ll >= 280) {
      tail += E;
      ++cnt;
    }
    if (ll % 50 == 0) std::printf("%d\t%.8f\t%.4e\t%.2e\t%ld\t%ld\n", ll, E, nv, cgErr, smp, slv);
  };
  runner.run(sampler, callback, randomizer, ham, 63, 200);
  const double e0 = -2.0 / std::sin(M_PI / (2 * N));
  tail /= cnt;
  std::printf("final %.8f %.8f\n", tail, e0);
  return std::fabs(tail - e0) / std::fabs(e0) < 0.005 ? 0 : 1;
}
