// Checkpoints through the concept layer: RunRBM with setIterParams(20, 10)
// writes w0000.dat, w0010.dat and w0020.dat into the working directory at the
// start of those epochs, as Runners/RunRBM.hpp:155-165 does; a second runner
// reads w0020.dat with initializeFrom (AbstractRunner.hpp:83-94) and must hold
// exactly the parameters the first had after 20 updates.  Exit status 0 on
// success.
#include <cstdio>
#include <sstream>

#include "yannq_amd/yannq_amd.hpp"

using namespace yannq_amd;

int main() {
  const int N = 20, alpha = 1;
  Device dev(0);
  std::ostringstream log;
  RunRBM<double> runner(dev, N, alpha, true, log);
  runner.setSeed(99);
  runner.initializeRandom(0.01);
  runner.setIterParams(20, 10);
  runner.setOptimizer("SGD", 0.01, 0.0);
  TFIsing ham(N, -1.0, -1.0);
  LocalSweeper sweeper(N);
  auto sampler = runner.createSampler(sweeper, 1, 16, 77);
  auto randomizer = [N](auto& re) { return randomSigma(N, re); };
  std::vector<double> after20;
  auto callback = [&](int ll, double, double, double, long, long) {
    if (ll == 19) after20 = runner.getQs().getParams();  // after the 20th update
  };
  runner.run(sampler, callback, randomizer, ham, 20, 20);
  for (const char* f : {"w0000.dat", "w0010.dat", "w0020.dat"}) {
    std::FILE* fp = std::fopen(f, "rb");
    if (!fp) {
      std::printf("missing %s\n", f);
      return 1;
    }
    std::fclose(fp);
  }
  Device dev2(0);
  std::ostringstream log2;
  RunRBM<double> second(dev2, N, alpha, true, log2);
  second.initializeFrom("w0020.dat");
  const auto p = second.getQs().getParams();
  if (p != after20) {
    std::printf("parameters differ after the round trip\n");
    return 1;
  }
  std::printf("checkpoints ok (%zu parameters)\n", p.size());
  return 0;
}
