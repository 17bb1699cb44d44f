// Header-only C++17 layer that rebuilds yannq's compile-time concepts --
// Machine (RBM), Hamiltonian, Sweeper, Sampler (SamplerMT), SR (SRMat),
// Optimizer and Runner (RunRBM) -- on top of the C ABI in yannq_amd.h, so
// code written against the reference's template API drops onto the GPU.
//
// Member names, argument meaning and defaults follow the reference
// (citations are paths under the yannq source tree).  Differences:
//   * std::vector replaces Eigen vectors (Eigen is not a dependency);
//   * C-ABI failures become yannq_amd::Error (a std::runtime_error) carrying
//     the yq_status code -- the reference mostly fails silently;
//   * random streams are counter-based Philox keyed by a seed, so
//     initializeRandomEngine() only re-keys the streams.
#pragma once

#include <array>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <fstream>
#include <functional>
#include <iterator>
#include <map>
#include <memory>
#include <ostream>
#include <random>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../yannq_amd.h"

namespace yannq_amd {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

namespace detail {
inline void check(yq_ctx* c, int rc, int allow = -1) {
  if (rc != YQ_OK && rc != allow)
    throw Error(rc, std::string("yannq_amd: ") + yq_last_error(c));
}
// Utilities/type_traits.hpp: is_complex_type / remove_complex
template <typename T>
struct is_complex : std::false_type {};
template <typename T>
struct is_complex<std::complex<T>> : std::true_type {};
template <typename T>
struct real_of {
  using type = T;
};
template <typename T>
struct real_of<std::complex<T>> {
  using type = T;
};
// the C ABI carries complex entries as interleaved (re, im) doubles
template <typename T>
double* dbl(T* p) {
  return reinterpret_cast<double*>(p);
}
template <typename T>
const double* dbl(const T* p) {
  return reinterpret_cast<const double*>(p);
}
}  // namespace detail

// One GPU context; every object below that lives on the device refers to one.
class Device {
 public:
  explicit Device(int device = 0) {
    const int rc = yq_ctx_create(device, &ctx_);
    if (rc != YQ_OK) throw Error(rc, "yannq_amd: cannot create a context on the GPU");
  }
  ~Device() { yq_ctx_destroy(ctx_); }
  Device(const Device&) = delete;
  Device& operator=(const Device&) = delete;
  yq_ctx* ctx() const { return ctx_; }

 private:
  yq_ctx* ctx_ = nullptr;
};

// ------------------------------------------------------------ Machine
// Machines/RBM.hpp:22-407 for T = double or std::complex<double>.
template <typename T = double>
class RBM {
  static_assert(std::is_same<T, double>::value || std::is_same<T, std::complex<double>>::value,
                "RBM<double> or RBM<std::complex<double>>");

 public:
  using Scalar = T;
  using RealScalar = typename detail::real_of<T>::type;
  using Vector = std::vector<T>;
  using DataT = std::tuple<std::vector<int>, Vector>;

  RBM(Device& dev, uint32_t n, uint32_t m, bool useBias = true)
      : dev_(&dev), n_(n), m_(m), useBias_(useBias) {
    const int dtype = detail::is_complex<T>::value ? YQ_DTYPE_C128 : YQ_DTYPE_F64;
    detail::check(ctx(), yq_rbm_create(ctx(), dtype, (int)n, (int)m, useBias ? 1 : 0));
  }
  yq_ctx* ctx() const { return dev_->ctx(); }

  uint32_t getN() const { return n_; }
  uint32_t getM() const { return m_; }
  uint32_t getDim() const { return useBias_ ? n_ * m_ + n_ + m_ : n_ * m_; }  // :124-130
  bool useBias() const { return useBias_; }
  std::string desc() const {  // :104-113
    return "{\"name\": \"RBM\", \"useBias\": " + std::string(useBias_ ? "true" : "false") +
           ", \"n\": " + std::to_string(n_) + ", \"m\": " + std::to_string(m_) + "}";
  }

  Vector getParams() const {  // :271-281
    Vector v(getDim());
    detail::check(ctx(), yq_rbm_get_params(ctx(), detail::dbl(v.data())));
    return v;
  }
  void setParams(const Vector& v) {  // :283-291
    if (v.size() != getDim()) throw Error(YQ_ERR_ARG, "setParams: wrong size");
    detail::check(ctx(), yq_rbm_set_params(ctx(), detail::dbl(v.data())));
  }
  void updateParams(const Vector& v) {  // :261-269
    if (v.size() != getDim()) throw Error(YQ_ERR_ARG, "updateParams: wrong size");
    detail::check(ctx(), yq_rbm_update_params(ctx(), detail::dbl(v.data())));
  }
  // :301-322 (complex: :327-349); the engine only supplies the Philox key
  template <class RandomEngine>
  void initializeRandom(RandomEngine& re, RealScalar sigma = 1e-3) {
    const uint64_t seed = ((uint64_t)re() << 32) ^ (uint64_t)re();
    detail::check(ctx(), yq_rbm_init_random(ctx(), seed, sigma));
  }
  bool hasNaN() const {  // :293-296
    int out = 0;
    detail::check(ctx(), yq_rbm_has_nan(ctx(), &out));
    return out != 0;
  }
  // The reference's weight file (cereal binary archive of RBM<T> through a
  // unique_ptr, Serializers/SerializeRBM.hpp:21-79), as RunRBM.hpp:155-165
  // writes w%04d.dat.
  void save(const std::string& path) const {
    detail::check(ctx(), yq_rbm_save(ctx(), path.c_str(), ckptElem(), 1));
  }
  // Read such a file into this machine (AbstractRunner.hpp:83-94); its N, M
  // and useBias must be this machine's.
  void load(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw Error(YQ_ERR_ARG, "cannot open " + path);
    const std::vector<unsigned char> buf((std::istreambuf_iterator<char>(in)),
                                         std::istreambuf_iterator<char>());
    int n = 0, m = 0, ub = 0;
    Vector v(getDim());
    const int rc = yq_ckpt_rbm_decode(ckptElem(), buf.data(), buf.size(), 1, &n, &m, &ub,
                                      nullptr, 0);
    if (rc != YQ_OK) throw Error(rc, path + ": not an RBM checkpoint of this scalar type");
    if ((uint32_t)n != n_ || (uint32_t)m != m_ || (ub != 0) != useBias_)
      throw Error(YQ_ERR_ARG, path + ": machine shape differs");
    const size_t len = v.size() * (detail::is_complex<T>::value ? 2 : 1);
    const int rc2 = yq_ckpt_rbm_decode(ckptElem(), buf.data(), buf.size(), 1, &n, &m, &ub,
                                       detail::dbl(v.data()), len);
    if (rc2 != YQ_OK) throw Error(rc2, path + ": malformed RBM checkpoint");
    setParams(v);
  }
  // W(j, i), A(i), B(j) readers (:205-216) through a parameter copy
  T W(uint32_t j, uint32_t i) const { return getParams()[(size_t)i * m_ + j]; }
  T A(uint32_t i) const { return useBias_ ? getParams()[(size_t)n_ * m_ + i] : T(0); }
  T B(uint32_t j) const { return useBias_ ? getParams()[(size_t)n_ * m_ + n_ + j] : T(0); }

 private:
  static int ckptElem() { return detail::is_complex<T>::value ? YQ_CKPT_C128 : YQ_CKPT_F64; }
  Device* dev_;
  uint32_t n_, m_;
  bool useBias_;
};

// ------------------------------------------------------------ Hamiltonians
// E_loc functors become descriptors; the ED column functors stay on the host.
class TFIsing {  // Hamiltonians/TFIsing.hpp
 public:
  TFIsing(int n, double J, double h) : n_(n), J_(J), h_(h) {}
  yq_ham_desc desc() const { return yq_ham_desc{YQ_TFI, n_, J_, h_, 0.0, 0.0, 0}; }
  std::map<uint32_t, double> operator()(uint32_t col) const {  // :53-65
    std::map<uint32_t, double> res;
    for (int i = 0; i < n_; i++) {
      const int s1 = (col >> i) & 1, s2 = (col >> ((i + 1) % n_)) & 1;
      res[col ^ (1u << i)] += h_;
      res[col] += J_ * (1 - 2 * s1) * (1 - 2 * s2);
    }
    return res;
  }
  std::string params() const {
    return "{\"name\": \"TFIsing\", \"n\": " + std::to_string(n_) + ", \"J\": " +
           std::to_string(J_) + ", \"h\": " + std::to_string(h_) + "}";
  }

 private:
  int n_;
  double J_, h_;
};

class XXZ {  // Hamiltonians/XXZ.hpp
 public:
  XXZ(int n, double J, double Delta, bool signRule = false)
      : n_(n), J_(J), Delta_(Delta), sign_(signRule ? -1.0 : 1.0) {}
  yq_ham_desc desc() const {
    return yq_ham_desc{YQ_XXZ, n_, J_, 0.0, Delta_, 0.0, sign_ < 0 ? 1 : 0};
  }
  std::map<uint32_t, double> operator()(uint32_t col) const {  // :51-64
    std::map<uint32_t, double> res;
    for (int i = 0; i < n_; i++) {
      const int b1 = (col >> i) & 1, b2 = (col >> ((i + 1) % n_)) & 1;
      const int zz = (1 - 2 * b1) * (1 - 2 * b2);
      res[col] += J_ * Delta_ * zz;
      res[col ^ ((1u << i) | (1u << ((i + 1) % n_)))] += sign_ * J_ * (1 - zz);
    }
    return res;
  }

 private:
  int n_;
  double J_, Delta_, sign_;
};

class J1J2 {  // Hamiltonians/J1J2.hpp
 public:
  J1J2(int n, double J1, double J2, bool signRule)
      : n_(n), J1_(J1), J2_(J2), sign_(signRule ? -1 : 1) {}
  yq_ham_desc desc() const {
    return yq_ham_desc{YQ_J1J2, n_, J1_, 0.0, 0.0, J2_, sign_ < 0 ? 1 : 0};
  }
  std::map<uint32_t, double> operator()(uint32_t col) const {  // :57-79
    std::map<uint32_t, double> res;
    for (int d = 1; d <= 2; ++d)
      for (int i = 0; i < n_; i++) {
        const int b1 = (col >> i) & 1, b2 = (col >> ((i + d) % n_)) & 1;
        const int zz = (1 - 2 * b1) * (1 - 2 * b2);
        const double Jd = d == 1 ? J1_ : J2_;
        res[col ^ ((1u << i) | (1u << ((i + d) % n_)))] +=
            Jd * (d == 1 ? sign_ : 1) * (1.0 - zz * 1.0);
        res[col] += Jd * zz;
      }
    return res;
  }

 private:
  int n_;
  double J1_, J2_;
  int sign_;
};

// ------------------------------------------------------------ Sweepers
class LocalSweeper {  // Samplers/LocalSweeper.hpp:11-54
 public:
  explicit LocalSweeper(uint32_t n, uint32_t nSweep = 1) : n_(n), nSweep_(nSweep) {}
  static constexpr int kind = YQ_LOCAL;
  uint32_t nSweep() const { return nSweep_; }
  std::string desc() const {
    return "{\"name\": \"Local Sweeper\", \"num_sweep_per\": " + std::to_string(nSweep_) + "}";
  }

 private:
  uint32_t n_, nSweep_;
};

class SwapSweeper {  // Samplers/SwapSweeper.hpp:8-57
 public:
  explicit SwapSweeper(uint32_t n, uint32_t nSweep = 1) : n_(n), nSweep_(nSweep) {}
  static constexpr int kind = YQ_EXCHANGE;
  uint32_t nSweep() const { return nSweep_; }
  std::string desc() const {
    return "{\"name\": \"Swap Sweeper\", \"num_sweeps_per\": " + std::to_string(nSweep_) + "}";
  }

 private:
  uint32_t n_, nSweep_;
};

// ------------------------------------------------------------ Sampler
// The batch of one sample() call; it stays on the device.  size() and
// operator[] (SamplerMT.hpp:195-203's container) copy out on first access.
template <class Machine>
class SampleBatch {
 public:
  using DataT = typename Machine::DataT;
  SampleBatch(const Machine& qs, int ns) : qs_(&qs), ns_(ns) {}
  std::size_t size() const { return (std::size_t)ns_; }
  DataT operator[](std::size_t idx) const {
    fetch();
    const int n = (int)qs_->getN(), m = (int)qs_->getM();
    std::vector<int> s(sig_.begin() + idx * n, sig_.begin() + (idx + 1) * n);
    typename Machine::Vector t(theta_.begin() + idx * m, theta_.begin() + (idx + 1) * m);
    return DataT(std::move(s), std::move(t));
  }

 private:
  void fetch() const {
    if (!sig_.empty()) return;
    sig8_.resize((size_t)ns_ * qs_->getN());
    theta_.resize((size_t)ns_ * qs_->getM());
    detail::check(qs_->ctx(), yq_samples_get(qs_->ctx(), sig8_.data(), detail::dbl(theta_.data())));
    sig_.assign(sig8_.begin(), sig8_.end());
  }
  const Machine* qs_;
  int ns_;
  mutable std::vector<int8_t> sig8_;
  mutable std::vector<int> sig_;
  mutable std::vector<typename Machine::Scalar> theta_;
};

// Samplers/SamplerMT.hpp:15-205 with nTmps == 1.
template <class Machine, class Sweeper>
class SamplerMT {
 public:
  using Scalar = typename Machine::Scalar;
  using RealScalar = typename Machine::RealScalar;

  SamplerMT(const Machine& qs, uint32_t nTmps, uint32_t nChainsPer, Sweeper& sweeper,
            uint64_t seed = 0x5eed, int precision = YQ_FP64)
      : qs_(qs), nTmps_(nTmps), nChainsPer_(nChainsPer), sweeper_(sweeper), seed_(seed),
        precision_(precision) {
    configure();  // nTmps: 1, or even and <= 16 (parallel tempering, :44-60, :120-153)
  }
  std::string desc() const {  // :62-70
    return "{\"name\": \"SamplerMT\", \"num_temps\": " + std::to_string(nTmps_) +
           ", \"num_chains_per_each_temp\": " + std::to_string(nChainsPer_) +
           ", \"sweeper\": " + sweeper_.desc() + "}";
  }
  // betas_ (:48-51): beta of state idx
  double beta(uint32_t idx) const { return double(nTmps_ - idx % nTmps_) / nTmps_; }
  // :185-192 -- re-keys the counter-based streams (keeps the seed by default)
  void initializeRandomEngine() { configure(); }
  void initializeRandomEngine(uint64_t seed) {
    seed_ = seed;
    configure();
  }
  void randomizeSigma() { detail::check(ctx(), yq_sampler_randomize(ctx(), -1)); }  // :194-202
  void randomizeSigma(int nup) { detail::check(ctx(), yq_sampler_randomize(ctx(), nup)); }
  // :216-223: randomizer(re) -> configuration, called once per chain state
  template <class Randomizer>
  void randomize(Randomizer&& randomizer) {
    std::mt19937_64 re(seed_ ^ (++epoch_ * 0x9E3779B97F4A7C15ull));
    const uint32_t n = qs_.getN();
    const uint32_t states = nTmps_ * nChainsPer_;
    std::vector<int8_t> all((size_t)states * n);
    for (uint32_t c = 0; c < states; ++c) {
      const auto s = randomizer(re);
      if ((uint32_t)s.size() != n) throw Error(YQ_ERR_ARG, "randomizer: wrong length");
      for (uint32_t i = 0; i < n; ++i) all[(size_t)c * n + i] = (int8_t)s[i];
    }
    detail::check(ctx(), yq_sampler_set_sigma(ctx(), all.data()));
  }
  // :293-317
  SampleBatch<Machine> sample(uint32_t nSweeps, uint32_t nDiscard) {
    int ns = 0;
    detail::check(ctx(), yq_sample(ctx(), (int)nSweeps, (int)nDiscard, &ns));
    return SampleBatch<Machine>(qs_, ns);
  }
  int precision() const { return precision_; }

 private:
  yq_ctx* ctx() const { return qs_.ctx(); }
  void configure() {
    detail::check(ctx(), yq_sampler_config(ctx(), Sweeper::kind, (int)nTmps_, (int)nChainsPer_,
                                           (int)sweeper_.nSweep(), precision_, seed_));
  }
  const Machine& qs_;
  uint32_t nTmps_, nChainsPer_;
  Sweeper& sweeper_;
  uint64_t seed_;
  int precision_;
  uint64_t epoch_ = 0;
};

// ------------------------------------------------------------ SR
// GroundState/SRMat.hpp:33-239.  The solver choice (explicit S + Cholesky,
// CG on the materialised O, or CG on the factored O) and the precision are
// fixed at construction.
template <class Machine, class Hamiltonian>
class SRMat {
 public:
  using Scalar = typename Machine::Scalar;
  using RealScalar = typename Machine::RealScalar;
  using Vector = std::vector<Scalar>;

  SRMat(const Machine& qs, const Hamiltonian& ham, int solver = YQ_SOLVE_EXPLICIT,
        int precision = YQ_FP64)
      : qs_(qs), ham_(ham), solver_(solver), precision_(precision) {}

  long rows() const { return (long)qs_.getDim(); }
  long cols() const { return (long)qs_.getDim(); }

  // :89-103
  template <class SamplingResult>
  void constructFromSamples(SamplingResult&& sr) {
    (void)sr;  // the batch of the last sample() call is on the device
    const yq_ham_desc d = ham_.desc();
    detail::check(ctx(), yq_ham_set(ctx(), &d));
    double mean = 0.0, var = 0.0;
    detail::check(ctx(), yq_eloc(ctx(), nullptr, &mean, &var));
    eloc_ = mean;
    elocVar_ = var;
    detail::check(ctx(), yq_sr_build(ctx(), solver_, precision_));
  }
  void setShift(RealScalar shift) { shift_ = shift; }
  RealScalar getShift() const { return shift_; }
  RealScalar eloc() const { return eloc_; }        // :151-154
  RealScalar elocVar() const { return elocVar_; }  // :156-159
  Vector energyGrad() const { return item(YQ_SR_F); }  // :161-168
  Vector oloc() const { return item(YQ_SR_OMEAN); }   // :132-140
  // :142-149, dim x dim column-major (explicit solver only)
  std::vector<Scalar> corrMat() const {
    const size_t d = qs_.getDim();
    std::vector<Scalar> S(d * d);
    detail::check(ctx(), yq_sr_get(ctx(), YQ_SR_S, detail::dbl(S.data())));
    return S;
  }
  Vector apply(const Vector& rhs) const {  // :170-189
    Vector out(rhs.size());
    detail::check(ctx(), yq_sr_apply(ctx(), shift_, detail::dbl(rhs.data()),
                                     detail::dbl(out.data())));
    return out;
  }
  // :197-223.  Unconverged CG returns its iterate, as Eigen's solve() does.
  Vector solveCG(RealScalar shift, RealScalar tol = 1e-4) { return solveCG(energyGrad(), shift, tol); }
  Vector solveCG(const Vector& vec, double shift, double tol = 1e-4) {
    setShift(shift);
    Vector x(vec.size());
    double res = 0.0;
    int it = 0;
    detail::check(ctx(), yq_sr_solve(ctx(), shift, tol, 2 * (int)vec.size(),
                                     detail::dbl(vec.data()), detail::dbl(x.data()), &res, &it),
                  YQ_ERR_NOT_CONVERGED);
    lastIters_ = it;
    lastResid_ = res;
    return x;
  }
  Vector solveExact(RealScalar shift) { return solveExact(energyGrad(), shift); }  // :231-237
  // RunRBM.hpp:194-203: Cholesky of S + shift I applied to an averaged gradient
  Vector solveExact(const Vector& rhs, RealScalar shift) {
    setShift(shift);
    Vector x(qs_.getDim());
    double res = 0.0;
    int it = 0;
    detail::check(ctx(), yq_sr_solve(ctx(), shift, 0.0, 0, detail::dbl(rhs.data()),
                                     detail::dbl(x.data()), &res, &it));
    lastResid_ = res;
    return x;
  }
  // fisherAvg.update(corrMat()) followed by getAvg() (RunRBM.hpp:13-36,
  // 197-198) on the device: S of this build becomes the bias-corrected
  // moving average that the next solveExact factors (explicit solver)
  void fisherAverage(double beta2, bool reset) {
    detail::check(ctx(), yq_sr_fisher_average(ctx(), beta2, reset ? 1 : 0));
  }
  int lastIterations() const { return lastIters_; }
  double lastResidual() const { return lastResid_; }
  int solver() const { return solver_; }

 private:
  yq_ctx* ctx() const { return qs_.ctx(); }
  Vector item(int which) const {
    Vector v(qs_.getDim());
    detail::check(ctx(), yq_sr_get(ctx(), which, detail::dbl(v.data())));
    return v;
  }
  const Machine& qs_;
  const Hamiltonian& ham_;
  int solver_, precision_;
  RealScalar shift_ = 0.0, eloc_ = 0.0, elocVar_ = 0.0;
  int lastIters_ = 0;
  double lastResid_ = 0.0;
};

// ------------------------------------------------------------ Optimizers
template <typename T>
class Optimizer {  // Optimizers/Optimizer.hpp:10-22
 public:
  using Vector = std::vector<T>;
  virtual std::string desc() const = 0;
  virtual Vector getUpdate(const Vector& grad) = 0;
  virtual ~Optimizer() {}
};

template <typename T>
class SGD : public Optimizer<T> {  // Optimizers/SGD.hpp
 public:
  using typename Optimizer<T>::Vector;
  explicit SGD(double alpha = 0.05, double p = 0.0) : alpha_(alpha), p_(p), t_(0) {}
  std::string desc() const override {
    return "{\"name\": \"SGD\", \"alhpa\": " + std::to_string(alpha_) + ", \"p\": " +
           std::to_string(p_) + "}";
  }
  double eta() const { return std::max(alpha_ / std::pow((double)t_, p_), 1e-4); }
  Vector getUpdate(const Vector& v) override {  // :48-54
    ++t_;
    const double e = eta();
    Vector out(v.size());
    for (size_t i = 0; i < v.size(); ++i) out[i] = -e * v[i];
    return out;
  }
  int step() const { return t_; }

 private:
  double alpha_, p_;
  int t_;
};

// Optimizers/OptimizerFactory.hpp:84-90: unknown names throw invalid_argument.
template <typename T>
std::unique_ptr<Optimizer<T>> createOptimizer(const std::string& name, double alpha = 0.05,
                                              double p = 0.0) {
  if (name == "SGD") return std::unique_ptr<Optimizer<T>>(new SGD<T>(alpha, p));
  throw std::invalid_argument("Such an optimizer does not exist.");
}

// ------------------------------------------------------------ exact sums
// Basis/BasisFull.hpp and Basis/BasisJz.hpp: which configurations a sum runs
// over (the device enumerates them in the same order).
struct BasisFull {
  explicit BasisFull(unsigned N) : N_(N) {}
  int nup() const { return -1; }
  unsigned N_;
};
struct BasisJz {
  BasisJz(unsigned N, unsigned nup) : N_(N), nup_(nup) {}
  int nup() const { return (int)nup_; }
  unsigned N_, nup_;
};

// GroundState/SRMatExact.hpp:47-149: expectation values as weighted sums
// over every basis configuration (|psi|^2 weights), on the device.
template <class Machine, class Hamiltonian>
class SRMatExact {
 public:
  using Scalar = typename Machine::Scalar;
  using RealScalar = typename Machine::RealScalar;
  using Vector = std::vector<Scalar>;

  template <class Basis>
  SRMatExact(const Machine& qs, Basis&& basis, const Hamiltonian& ham)
      : qs_(qs), ham_(ham), nup_(basis.nup()) {}

  void clear() { eloc_ = elocVar_ = 0.0; }
  void constructExact() {  // :97-117
    const yq_ham_desc d = ham_.desc();
    detail::check(ctx(), yq_ham_set(ctx(), &d));
    int B = 0;
    detail::check(ctx(), yq_exact_batch(ctx(), nup_, &B));
    double mean = 0.0, var = 0.0;
    detail::check(ctx(), yq_eloc(ctx(), nullptr, &mean, &var));
    eloc_ = mean;
    elocVar_ = var;
    detail::check(ctx(), yq_sr_build(ctx(), YQ_SOLVE_EXPLICIT, YQ_FP64));
  }
  RealScalar eloc() const { return eloc_; }
  RealScalar elocVar() const { return elocVar_; }
  Vector oloc() const { return item(YQ_SR_OMEAN); }      // :119-127
  Vector energyGrad() const { return item(YQ_SR_F); }    // :136-144
  std::vector<Scalar> corrMat() const {                   // :129-134, column-major
    const size_t d = qs_.getDim();
    std::vector<Scalar> S(d * d);
    detail::check(ctx(), yq_sr_get(ctx(), YQ_SR_S, detail::dbl(S.data())));
    return S;
  }
  Vector apply(const Vector& rhs) const {  // :146-151
    Vector out(rhs.size());
    detail::check(ctx(), yq_sr_apply(ctx(), 0.0, detail::dbl(rhs.data()),
                                     detail::dbl(out.data())));
    return out;
  }
  // RunRBMExact.hpp:57-66: LLT of corrMat + shift I applied to rhs
  Vector solveExact(const Vector& rhs, RealScalar shift) const {
    Vector x(qs_.getDim());
    double res = 0.0;
    int it = 0;
    detail::check(ctx(), yq_sr_solve(ctx(), shift, 0.0, 0, detail::dbl(rhs.data()),
                                     detail::dbl(x.data()), &res, &it));
    return x;
  }

 private:
  yq_ctx* ctx() const { return qs_.ctx(); }
  Vector item(int which) const {
    Vector v(qs_.getDim());
    detail::check(ctx(), yq_sr_get(ctx(), which, detail::dbl(v.data())));
    return v;
  }
  const Machine& qs_;
  const Hamiltonian& ham_;
  int nup_;
  RealScalar eloc_ = 0.0, elocVar_ = 0.0;
};

// Samplers/ExactSampler.hpp:11-79: independent draws from |psi|^2
template <class Machine>
class ExactSampler {
 public:
  template <class Basis>
  ExactSampler(const Machine& qs, Basis&& basis) : qs_(qs), nup_(basis.nup()) {}
  void initializeRandomEngine() { seed_ = std::random_device{}(); }
  void setSeed(uint64_t s) { seed_ = s; }
  template <class Randomizer>
  void randomize(Randomizer&&) const {}  // nothing to randomise
  // sample(n_sweeps, .): n_sweeps draws; returns their basis positions (the
  // configurations are the device batch, as SamplerMT's are)
  std::vector<int> sample(uint32_t nSweeps, uint32_t /*nDiscard*/ = 0) {
    std::vector<int> idx(nSweeps);
    detail::check(qs_.ctx(), yq_exact_sample(qs_.ctx(), nup_, (int)nSweeps, seed_++, idx.data()));
    return idx;
  }
  std::string desc() const { return "{\"name\": \"ExactSampler\"}"; }

 private:
  const Machine& qs_;
  int nup_;
  uint64_t seed_ = 0x5eed;
};

// ------------------------------------------------------------ Runner
// Runners/RunRBM.hpp:41-222 with AbstractRunner.hpp's defaults (:32-38).
template <typename T = double>
class RunRBM {
 public:
  using Machine = RBM<T>;
  using Vector = std::vector<T>;

  RunRBM(Device& dev, unsigned N, unsigned alpha, bool useBias, std::ostream& logger)
      : logger_(logger), qs_(dev, N, alpha * N, useBias), re_(std::random_device{}()) {}

  void setLambda(double ini, double decay = 1.0, double min = 0.0) {
    lambdaIni_ = ini;
    lambdaDecay_ = decay;
    lambdaMin_ = min;
  }
  void setIterParams(int maxIter, int saveWfPer) {
    maxIter_ = maxIter;
    saveWfPer_ = saveWfPer;
  }
  void setSolverParams(bool useCG, double cgTol = 1e-4) {
    useCG_ = useCG;
    cgTol_ = cgTol;
  }
  // The first-moment averaging (beta1) runs on the host; the fisher average
  // (beta2) on the device, on the explicit S of every epoch.
  void setMomentum(double beta1 = 0.0, double beta2 = 0.0) {
    beta1_ = beta1;
    beta2_ = beta2;
  }
  void setGradClip(double c) { gradClip_ = c; }
  void setOptimizer(const std::string& name, double alpha = 0.05, double p = 0.0) {
    opt_ = createOptimizer<T>(name, alpha, p);
  }
  void setSeed(uint64_t s) { re_.seed(s); }
  void initializeFrom(const std::string& filePath) {  // AbstractRunner.hpp:83-94
    machineInitialized_ = true;
    logger_ << "Loading initial weights from " << filePath << std::endl;
    qs_.load(filePath);
  }
  void initializeRandom(double wIni = 1e-3) {  // AbstractRunner.hpp:96-101
    machineInitialized_ = true;
    logger_ << "Set initial weights randomly from N(0.0, " << wIni << "^2)" << std::endl;
    qs_.initializeRandom(re_, wIni);
  }
  Machine& getQs() { return qs_; }
  uint32_t getDim() const { return qs_.getDim(); }
  template <class Sweeper>
  SamplerMT<Machine, Sweeper> createSampler(Sweeper& sweeper, uint32_t nTmps,
                                            uint32_t nChainsPer, uint64_t seed = 0x5eed,
                                            int precision = YQ_FP64) const {
    return SamplerMT<Machine, Sweeper>(qs_, nTmps, nChainsPer, sweeper, seed, precision);
  }

  // :122-221.  callback(ll, E, |v|, cgErr, smp_ms, slv_ms).  The solver is
  // explicit S + Cholesky unless setSolverParams(true, .) with beta2 = 0,
  // which selects CG on the factored O (precision given by the sampler).
  template <class Sampler, class Callback, class Randomizer, class Hamiltonian>
  void run(Sampler& sampler, Callback&& callback, Randomizer&& randomizer, Hamiltonian&& ham,
           int nSweeps = -1, int nSweepsDiscard = -1) {
    using Clock = std::chrono::high_resolution_clock;
    if (!machineInitialized_) initializeRandom();
    if (!opt_) setOptimizer("SGD");
    const bool cg = useCG_ && beta2_ == 0.0;  // RunRBM.hpp:189
    const int solver = cg ? YQ_SOLVE_CG_FACTORED : YQ_SOLVE_EXPLICIT;
    const int precision = cg ? sampler.precision() : YQ_FP64;
    SRMat<Machine, typename std::decay<Hamiltonian>::type> srm(qs_, ham, solver, precision);
    sampler.initializeRandomEngine();
    const int dim = (int)getDim();
    if (nSweeps == -1) nSweeps = dim;
    if (nSweepsDiscard == -1) nSweepsDiscard = int(0.1 * nSweeps);
    Vector gradAvg(dim, T(0.0));
    int t = 0;
    for (int ll = 0; ll <= maxIter_; ll++) {
      logger_ << "Epochs: " << ll << "\t# of discard samples: " << nSweepsDiscard << std::endl;
      if (saveWfPer_ != 0 && ll % saveWfPer_ == 0) {  // RunRBM.hpp:155-165
        char fileName[30];
        snprintf(fileName, sizeof(fileName), "w%04d.dat", ll);
        qs_.save(fileName);
      }
      sampler.randomize(randomizer);
      const auto smp_start = Clock::now();
      auto sr = sampler.sample(nSweeps, nSweepsDiscard);
      const auto smp_dur = std::chrono::duration_cast<std::chrono::milliseconds>(
                               Clock::now() - smp_start).count();
      logger_ << "Number of samples: " << sr.size() << std::endl;
      const auto slv_start = Clock::now();
      srm.constructFromSamples(sr);
      const double currE = srm.eloc();
      const double lambda = std::max(lambdaIni_ * std::pow(lambdaDecay_, ll), lambdaMin_);
      // ObsAvg with beta1 (RunRBM.hpp:13-36): bias-corrected moving average
      const Vector g = srm.energyGrad();
      auto grad_update = [&]() {
        ++t;
        for (int i = 0; i < dim; ++i) gradAvg[i] = beta1_ * gradAvg[i] + (1.0 - beta1_) * g[i];
      };
      grad_update();
      Vector grad(dim);
      for (int i = 0; i < dim; ++i) grad[i] = gradAvg[i] / (1.0 - std::pow(beta1_, t));
      Vector v;
      double cgErr = 0.0;
      if (cg) {
        v = srm.solveCG(grad, lambda, cgTol_);
        cgErr = srm.lastResidual() * norm(grad);  // |apply(v) - grad|
      } else {
        // :196-202: the gradient average is updated a second time (the
        // solve uses the first), S enters fisherAvg (identity at beta2 = 0)
        grad_update();
        if (beta2_ != 0.0) srm.fisherAverage(beta2_, ll == 0);
        v = srm.solveExact(grad, lambda);
      }
      const double nv = norm(v);
      if (gradClip_ != 0.0 && nv > gradClip_)
        for (auto& x : v) x *= gradClip_ / nv;
      const Vector optV = opt_->getUpdate(v);
      const auto slv_dur = std::chrono::duration_cast<std::chrono::milliseconds>(
                               Clock::now() - slv_start).count();
      qs_.updateParams(optV);
      callback(ll, currE, nv, cgErr, smp_dur, slv_dur);
    }
  }

 private:
  static double norm(const Vector& v) {
    double s = 0.0;
    for (const T& x : v) s += std::norm(x);
    return std::sqrt(s);
  }
  std::ostream& logger_;
  Machine qs_;
  std::mt19937_64 re_;
  std::unique_ptr<Optimizer<T>> opt_;
  bool machineInitialized_ = false, useCG_ = false;
  double cgTol_ = 1e-4, beta1_ = 0.0, beta2_ = 0.0, gradClip_ = 0.0;
  double lambdaIni_ = 1e-3, lambdaDecay_ = 1.0, lambdaMin_ = 1e-4;
  int maxIter_ = 2000, saveWfPer_ = 100;
};

// Runners/RunRBMExact.hpp:8-82: SR with exact sums over a basis,
// corrMat + lambda I by Cholesky, then the optimizer; callback(ll, E, |v|).
template <typename T = double>
class RunRBMExact {
 public:
  using Machine = RBM<T>;
  using Vector = std::vector<T>;

  RunRBMExact(Device& dev, unsigned N, unsigned alpha, bool useBias, std::ostream& logger)
      : logger_(logger), qs_(dev, N, alpha * N, useBias), re_(std::random_device{}()) {}
  void setLambda(double ini, double decay = 1.0, double min = 0.0) {
    lambdaIni_ = ini;
    lambdaDecay_ = decay;
    lambdaMin_ = min;
  }
  void setIterParams(int maxIter, int saveWfPer) {
    maxIter_ = maxIter;
    saveWfPer_ = saveWfPer;
  }
  void setOptimizer(const std::string& name, double alpha = 0.05, double p = 0.0) {
    opt_ = createOptimizer<T>(name, alpha, p);
  }
  void setSeed(uint64_t s) { re_.seed(s); }
  void initializeRandom(double wIni = 1e-3) {
    machineInitialized_ = true;
    qs_.initializeRandom(re_, wIni);
  }
  void initializeFrom(const std::string& filePath) {
    machineInitialized_ = true;
    qs_.load(filePath);
  }
  Machine& getQs() { return qs_; }

  template <class Callback, class Basis, class Hamiltonian>
  void run(Callback&& callback, Basis&& basis, Hamiltonian&& ham) {
    if (!machineInitialized_) initializeRandom();
    if (!opt_) setOptimizer("SGD");
    SRMatExact<Machine, typename std::decay<Hamiltonian>::type> srex(qs_, basis, ham);
    for (int ll = 0; ll <= maxIter_; ll++) {
      logger_ << "Epochs: " << ll << std::endl;
      if (saveWfPer_ != 0 && ll % saveWfPer_ == 0) {
        char fileName[30];
        snprintf(fileName, sizeof(fileName), "w%04d.dat", ll);
        qs_.save(fileName);
      }
      srex.clear();
      srex.constructExact();
      const double currE = srex.eloc();
      const double lambda = std::max(lambdaIni_ * std::pow(lambdaDecay_, ll), lambdaMin_);
      const Vector v = srex.solveExact(srex.energyGrad(), lambda);
      const Vector optV = opt_->getUpdate(v);
      qs_.updateParams(optV);
      double nv = 0.0;
      for (const T& x : v) nv += std::norm(x);
      callback(ll, currE, std::sqrt(nv));
    }
  }

 private:
  std::ostream& logger_;
  Machine qs_;
  std::mt19937_64 re_;
  std::unique_ptr<Optimizer<T>> opt_;
  bool machineInitialized_ = false;
  double lambdaIni_ = 1e-3, lambdaDecay_ = 1.0, lambdaMin_ = 1e-4;
  int maxIter_ = 2000, saveWfPer_ = 100;
};

// ------------------------------------------------------------ FeedForward
// Machines/FeedForward.hpp:16-205: any sequence of addLayer<L>(args...) with
// the reference's layer constructor arguments (layers/Conv1D.hpp:56-64,
// layers/FullyConnected.hpp:57-58, layers/Mean.hpp:21, ActivationLayer.hpp:
// 138-172 with activations.hpp's constructors).  The layers are kept on the
// host; the device network is made for an input of N spins by bind(N), or
// by the first forward / backward (N = the configuration's length) or the
// AmplitudePhase that takes the network.  Parameters set or drawn before
// that are kept and uploaded then.  Scalar is double.
enum class InitializationMode { LeCun, Xavier, He };  // FeedForward.hpp:11-14

struct Conv1D {
  Conv1D(uint32_t inChannels, uint32_t outChannels, uint32_t kernelSize, uint32_t stride = 1,
         bool useBias = false)
      : in(inChannels), out(outChannels), kernel(kernelSize), stride(stride), useBias(useBias) {}
  uint32_t in, out, kernel, stride;
  bool useBias;
};
struct FullyConnected {
  FullyConnected(uint32_t inputDim, uint32_t outputDim, bool useBias = false)
      : in(inputDim), out(outputDim), useBias(useBias) {}
  uint32_t in, out;
  bool useBias;
};
struct Mean {
  Mean(const std::array<uint32_t, 2>& inputDim, uint32_t meanAxis)
      : dim(inputDim), axis(meanAxis) {}
  std::array<uint32_t, 2> dim;
  uint32_t axis;
};
// activation layers (ActivationLayer.hpp:138-172); the constructor argument
// and its default are activations.hpp's
namespace detail {
template <int A>
struct Activation {
  explicit Activation(double p = 0.0) : param(p) {}
  double param;
};
}  // namespace detail
using Identity = detail::Activation<YQ_ACT_IDENTITY>;
using LnCosh = detail::Activation<YQ_ACT_LNCOSH>;
struct Tanh : detail::Activation<YQ_ACT_TANH> {  // constant * tanh(z / constant)
  explicit Tanh(double constant = 1.0) : Activation(constant) {}
};
using WeakTanh = detail::Activation<YQ_ACT_WEAKTANH>;
using Sigmoid = detail::Activation<YQ_ACT_SIGMOID>;
using ReLU = detail::Activation<YQ_ACT_RELU>;
struct LeakyReLU : detail::Activation<YQ_ACT_LEAKYRELU> {
  explicit LeakyReLU(double negativeSlope = 0.1) : Activation(negativeSlope) {}
};
using HardTanh = detail::Activation<YQ_ACT_HARDTANH>;
using SoftShrink = detail::Activation<YQ_ACT_SOFTSHRINK>;
struct LeakyHardTanh : detail::Activation<YQ_ACT_LEAKYHARDTANH> {
  explicit LeakyHardTanh(double outsideSlope = 0.01) : Activation(outsideSlope) {}
};
using SoftSign = detail::Activation<YQ_ACT_SOFTSIGN>;
struct Cos : detail::Activation<YQ_ACT_COS> {  // a cos(pi z / a)
  explicit Cos(double a = 1.0) : Activation(a) {}
};

class FeedForward {
 public:
  using Scalar = double;
  using Vector = std::vector<double>;

  explicit FeedForward(Device& dev) : dev_(&dev) {}
  FeedForward(FeedForward&&) = default;
  yq_ctx* ctx() const { return dev_->ctx(); }
  Device& device() const { return *dev_; }

  template <class Layer, class... Ts>
  void addLayer(Ts&&... args) {  // :76-82
    push(Layer(std::forward<Ts>(args)...));
    if (n_) bind(n_);  // the device network follows the layer list
  }
  bool complete() const { return !layers_.empty(); }
  uint32_t getN() const { return n_; }
  uint32_t getDim() const {  // :84-91 (the sum of the layers' paramDim)
    long long np = 0;
    for (const yq_ff_layer& l : layers_) np += paramDim(l);
    return (uint32_t)np;
  }
  // the device network for an input of N spins (made again when N changes)
  void bind(uint32_t N) {
    if (layers_.empty()) throw Error(YQ_ERR_STATE, "FeedForward: no layers");
    Vector keep = bound() ? getParams() : pending_;
    detail::check(ctx(), yq_ff_create_stack(ctx(), (int)N, 1, (int)layers_.size(),
                                            layers_.data()));
    n_ = N;
    if (keep.size() == getDim()) detail::check(ctx(), yq_ff_set_params(ctx(), keep.data()));
    pending_.clear();
  }
  std::string summary() const {  // :101-116
    std::string s = "Layer name\tNumber of params\n";
    for (const yq_ff_layer& l : layers_)
      s += std::string(layerName(l)) + "\t" + std::to_string(paramDim(l)) + "\n";
    return s;
  }
  std::string desc() const {  // :118-126, the layers' desc() (Conv1D.hpp:236-246 etc.)
    std::string s = "[";
    for (size_t i = 0; i < layers_.size(); ++i) {
      const yq_ff_layer& l = layers_[i];
      s += i ? ", {" : "{";
      s += "\"name\": \"" + std::string(layerName(l)) + "\"";
      if (l.kind == YQ_LAYER_CONV1D)
        s += ", \"use_bias\": " + std::string(l.use_bias ? "true" : "false") +
             ", \"input_channels\": " + std::to_string(l.a) + ", \"output_channels\": " +
             std::to_string(l.b) + ", \"kernel_size\": " + std::to_string(l.c) +
             ", \"stride\": " + std::to_string(l.d);
      else if (l.kind == YQ_LAYER_FC)
        s += ", \"use_bias\": " + std::string(l.use_bias ? "true" : "false") +
             ", \"input_dim\": " + std::to_string(l.a) + ", \"output_dim\": " +
             std::to_string(l.b);
      else if (l.kind == YQ_LAYER_MEAN)
        s += ", \"input_dim\": [" + std::to_string(l.a) + ", " + std::to_string(l.b) +
             "], \"mean_axis\": " + std::to_string(l.c);
      else
        s += ", \"activation\": \"" + std::string(activationName(l.a)) + "\"";
      s += "}";
    }
    return s + "]";
  }
  Vector getParams() const {
    if (!bound()) return pending_.empty() ? Vector(getDim(), 0.0) : pending_;
    Vector v(getDim());
    detail::check(ctx(), yq_ff_get_params(ctx(), v.data()));
    return v;
  }
  void setParams(const Vector& v) {
    if (v.size() != getDim()) throw Error(YQ_ERR_ARG, "setParams: wrong size");
    if (!bound()) {
      pending_ = v;
      return;
    }
    detail::check(ctx(), yq_ff_set_params(ctx(), v.data()));
  }
  void updateParams(const Vector& v) {
    if (v.size() != getDim()) throw Error(YQ_ERR_ARG, "updateParams: wrong size");
    if (!bound()) {
      Vector p = getParams();
      for (size_t i = 0; i < p.size(); ++i) p[i] += v[i];
      pending_ = p;
      return;
    }
    detail::check(ctx(), yq_ff_update_params(ctx(), v.data()));
  }
  // :43-73: layer by layer, paramDim draws of N(0, sigma(fanIn, fanOut)^2)
  // from the engine (randomVector, Utilities/Utility.hpp:126-138)
  template <class RandomEngine>
  void initializeRandom(RandomEngine& re, InitializationMode mode) {
    Vector v;
    v.reserve(getDim());
    for (const yq_ff_layer& l : layers_) {
      if (paramDim(l) == 0) continue;
      const double fi = l.kind == YQ_LAYER_CONV1D ? (double)l.a * l.c : (double)l.a;
      const double fo = l.kind == YQ_LAYER_CONV1D ? (double)l.b * l.c : (double)l.b;
      const double sigma = mode == InitializationMode::LeCun   ? std::sqrt(1.0 / fi)
                           : mode == InitializationMode::Xavier ? std::sqrt(2.0 / (fi + fo))
                                                                : std::sqrt(2.0 / fi);
      std::normal_distribution<double> nd(0.0, sigma);
      for (long long k = 0; k < paramDim(l); ++k) v.push_back(nd(re));
    }
    setParams(v);
  }
  // :33-41: every parameter from N(0, sigma^2), drawn with the engine
  template <class RandomEngine>
  void initializeRandom(RandomEngine& re, double sigma) {
    std::normal_distribution<double> nd(0.0, sigma);
    Vector v(getDim());
    for (auto& x : v) x = nd(re);
    setParams(v);
  }
  // forward(sigma) (:152-168) and backward(makeData(sigma)) (:170-205)
  double forward(const std::vector<int>& sigma) {
    const std::vector<int8_t> s = spins(sigma);
    double out = 0.0;
    detail::check(ctx(), yq_ff_forward(ctx(), 1, s.data(), &out));
    return out;
  }
  Vector backward(const std::vector<int>& sigma) {
    const std::vector<int8_t> s = spins(sigma);
    Vector der(getDim());
    detail::check(ctx(), yq_ff_jacobian(ctx(), 1, s.data(), der.data(), nullptr));
    return der;
  }

 private:
  bool bound() const { return n_ != 0; }
  static long long paramDim(const yq_ff_layer& l) {
    if (l.kind == YQ_LAYER_CONV1D) return (long long)l.a * l.b * l.c + (l.use_bias ? l.b : 0);
    if (l.kind == YQ_LAYER_FC) return (long long)l.a * l.b + (l.use_bias ? l.b : 0);
    return 0;
  }
  static const char* layerName(const yq_ff_layer& l) {
    switch (l.kind) {
      case YQ_LAYER_CONV1D: return "Convolutional 1D Layer";
      case YQ_LAYER_FC: return "Fully Connected Layer";
      case YQ_LAYER_MEAN: return "Mean layer";
      default: return "Activation Layer";
    }
  }
  static const char* activationName(int a) {
    static const char* const names[] = {"Identity", "LnCosh", "Tanh", "WeakTanh",
                                        "Sigmoid", "ReLU", "LeakyReLU", "HardTanh",
                                        "SoftShrink", "LeakyHardTanh", "SoftSign", "Cos"};
    return a >= 0 && a < 12 ? names[a] : "?";
  }
  std::vector<int8_t> spins(const std::vector<int>& sigma) {
    if (sigma.empty()) throw Error(YQ_ERR_ARG, "FeedForward: empty input");
    if (sigma.size() != n_) bind((uint32_t)sigma.size());
    return std::vector<int8_t>(sigma.begin(), sigma.end());
  }
  void add(int kind, int a, int b, int c, int d, bool bias, double param) {
    yq_ff_layer l{};
    l.kind = kind, l.a = a, l.b = b, l.c = c, l.d = d, l.use_bias = bias ? 1 : 0;
    l.param = param;
    layers_.push_back(l);
    pending_.clear();
  }
  void push(const Conv1D& l) {
    add(YQ_LAYER_CONV1D, (int)l.in, (int)l.out, (int)l.kernel, (int)l.stride, l.useBias, 0.0);
  }
  void push(const FullyConnected& l) {
    add(YQ_LAYER_FC, (int)l.in, (int)l.out, 0, 0, l.useBias, 0.0);
  }
  void push(const Mean& l) {
    add(YQ_LAYER_MEAN, (int)l.dim[0], (int)l.dim[1], (int)l.axis, 0, false, 0.0);
  }
  template <int A>
  void push(const detail::Activation<A>& l) {
    add(YQ_LAYER_ACT, A, 0, 0, 0, false, l.param);
  }
  Device* dev_;
  std::vector<yq_ff_layer> layers_;
  Vector pending_;  // parameters given before the device network exists
  uint32_t n_ = 0;
};

// ------------------------------------------------------------ AmplitudePhase
// Machines/AmplitudePhase.hpp:9-178: psi = sqrt(RBM) exp(i pi FF) with a
// real RBM(N, M, useBias) as the amplitude and the FeedForward network (on
// the same Device) as the phase.  The device machine samples, measures and
// solves for psi directly (yq_ap_enable); parameters are [RBM | FF].
class AmplitudePhase {
 public:
  using Scalar = double;
  using RealScalar = double;
  using ComplexScalar = std::complex<double>;
  using Vector = std::vector<double>;
  using AmplitudeMachine = RBM<double>;

  AmplitudePhase(Device& dev, uint32_t N, uint32_t M, bool useBias, FeedForward&& phase)
      : N_(N), amplitude_(dev, N, M, useBias), phase_(std::move(phase)) {  // :37-40
    if (&phase_.device() != &dev) throw Error(YQ_ERR_ARG, "AmplitudePhase: phase on another Device");
    if (!phase_.complete()) throw Error(YQ_ERR_ARG, "AmplitudePhase: the phase network is empty");
    phase_.bind(N);  // the phase network acts on the N spins
    detail::check(ctx(), yq_ap_enable(ctx(), 1));
  }
  yq_ctx* ctx() const { return amplitude_.ctx(); }

  uint32_t getN() const { return N_; }
  uint32_t getDimAmp() const { return amplitude_.getDim(); }  // :60-63
  uint32_t getDimPhase() const { return phase_.getDim(); }    // :65-68
  uint32_t getDim() const { return getDimAmp() + getDimPhase(); }
  const AmplitudeMachine& amplitudeMachine() const { return amplitude_; }
  const FeedForward& phaseMachine() const { return phase_; }

  template <class RandomEngine>
  void initializeAmplitudeRandom(RandomEngine& re, double sigma) {  // :85-89
    amplitude_.initializeRandom(re, sigma);
  }
  template <class RandomEngine>
  void initializeRandom(RandomEngine& re, double sigma) {  // :91-96
    initializeAmplitudeRandom(re, sigma);
    phase_.initializeRandom(re, InitializationMode::Xavier);
  }
  template <class RandomEngine>
  void initializePhase(RandomEngine& re, InitializationMode mode = InitializationMode::LeCun) {
    phase_.initializeRandom(re, mode);  // :99-103
  }
  Vector getParams() const {
    Vector v(getDim());
    detail::check(ctx(), yq_ap_get_params(ctx(), v.data()));
    return v;
  }
  void updateParams(const Vector& m) {  // :153-158
    if (m.size() != getDim()) throw Error(YQ_ERR_ARG, "updateParams: wrong size");
    detail::check(ctx(), yq_ap_update_params(ctx(), m.data()));
  }
  // log coeff(sigma) = log RBM(sigma) / 2 + i pi FF(sigma); coeff (:127-138)
  ComplexScalar logCoeff(const std::vector<int>& sigma) const {
    if (sigma.size() != N_) throw Error(YQ_ERR_ARG, "logCoeff: wrong input size");
    const std::vector<int8_t> s(sigma.begin(), sigma.end());
    double out[2];
    detail::check(ctx(), yq_ap_log_psi(ctx(), 1, s.data(), out));
    return ComplexScalar(out[0], out[1]);
  }
  ComplexScalar coeff(const std::vector<int>& sigma) const { return std::exp(logCoeff(sigma)); }
  double phaseForward(const std::vector<int>& sigma) { return phase_.forward(sigma); }
  std::string desc() const {  // :160-166
    return "{\"amplitude\": " + amplitude_.desc() + ", \"phase\": " + phase_.desc() + "}";
  }

 private:
  uint32_t N_;
  AmplitudeMachine amplitude_;
  FeedForward phase_;
};

// GroundState/NGDExact.hpp:17-204: exact energy, gradient and the amplitude
// and phase blocks of the metric over a basis, from the device's weighted
// basis batch.  energyGrad() keeps the reference's factor 2 (it is twice
// the real part of SRMatExact's gradient).
class NGDExact {
 public:
  using RealScalar = double;
  using RealVector = std::vector<double>;
  using RealMatrix = std::vector<double>;  // column-major

  template <class Basis, class Hamiltonian>
  NGDExact(const AmplitudePhase& qs, Basis&& basis, const Hamiltonian& ham)
      : qs_(qs), nup_(basis.nup()), ham_(ham.desc()) {}

  void constructExact() {  // :137-170
    detail::check(ctx(), yq_ham_set(ctx(), &ham_));
    int B = 0;
    detail::check(ctx(), yq_exact_batch(ctx(), nup_, &B));
    double var = 0.0;
    detail::check(ctx(), yq_eloc(ctx(), nullptr, &energy_, &var));
    detail::check(ctx(), yq_sr_build(ctx(), YQ_SOLVE_EXPLICIT, YQ_FP64));
    const size_t D = qs_.getDim();
    F_.assign(D, 0.0);
    oloc_.assign(D, 0.0);
    S_.assign(D * D, 0.0);
    detail::check(ctx(), yq_sr_get(ctx(), YQ_SR_F, F_.data()));
    detail::check(ctx(), yq_sr_get(ctx(), YQ_SR_OMEAN, oloc_.data()));
    detail::check(ctx(), yq_sr_get(ctx(), YQ_SR_S, S_.data()));
  }
  RealScalar eloc() const { return energy_; }
  RealScalar getEnergy() const { return energy_; }
  RealVector olocAmp() const { return RealVector(oloc_.begin(), oloc_.begin() + dA()); }
  RealVector olocPhase() const { return RealVector(oloc_.begin() + dA(), oloc_.end()); }
  RealMatrix corrMatAmp() const { return block(0, dA()); }                      // :172-177
  RealMatrix corrMatPhase() const { return block(dA(), qs_.getDim() - dA()); }  // :179-184
  RealVector energyGrad() const {
    RealVector g(F_);
    for (auto& x : g) x *= 2.0;
    return g;
  }
  // (block + lambda I) x = rhs for the amplitude (off 0) or phase block on
  // the device; rhs and x span the whole parameter vector, the other block's
  // entries of x are the solve of that block (the blocks are independent)
  RealVector solve(const RealVector& rhs, double lambda) const {
    RealVector x(rhs.size());
    double res = 0.0;
    int it = 0;
    detail::check(ctx(), yq_sr_solve(ctx(), lambda, 0.0, 0, rhs.data(), x.data(), &res, &it));
    return x;
  }

 private:
  yq_ctx* ctx() const { return qs_.ctx(); }
  size_t dA() const { return qs_.getDimAmp(); }
  RealMatrix block(size_t off, size_t d) const {
    const size_t D = qs_.getDim();
    RealMatrix out(d * d);
    for (size_t j = 0; j < d; ++j)
      for (size_t i = 0; i < d; ++i) out[j * d + i] = S_[(off + j) * D + off + i];
    return out;
  }
  const AmplitudePhase& qs_;
  int nup_;
  yq_ham_desc ham_;
  double energy_ = 0.0;
  RealVector F_, oloc_, S_;
};

// Runners/RunAmplitudePhase.hpp:13-122: NGDExact epochs with the amplitude
// and phase blocks solved separately (each + lambda I), then the optimizer;
// callback(ll, E, |grad|, |v|).  Weight files are not written (saveWfPer
// must be 0: the reference's AmplitudePhase archive holds the phase
// network's layer objects).
enum class Mode { AMP, PHASE, ALL };

class RunAmplitudePhaseExact {
 public:
  using Machine = AmplitudePhase;
  using RealVector = std::vector<double>;

  RunAmplitudePhaseExact(Device& dev, uint32_t N, uint32_t alpha, bool useBias,
                         FeedForward&& phase, std::ostream& logger)
      : logger_(logger), qs_(dev, N, N * alpha, useBias, std::move(phase)),
        re_(std::random_device{}()) {}
  void setLambda(double ini, double decay = 1.0, double min = 0.0) {
    lambdaIni_ = ini;
    lambdaDecay_ = decay;
    lambdaMin_ = min;
  }
  void setIterParams(int maxIter, int saveWfPer) {
    if (saveWfPer != 0)
      throw Error(YQ_ERR_UNSUPPORTED, "RunAmplitudePhaseExact: weight files are not written");
    maxIter_ = maxIter;
  }
  void setOptimizer(const std::string& name, double alpha = 0.05, double p = 0.0) {
    opt_ = createOptimizer<double>(name, alpha, p);
  }
  void setSeed(uint64_t s) { re_.seed(s); }
  void initializePhase() { qs_.initializePhase(re_, InitializationMode::Xavier); }  // :40-43
  void initializeAmplitude(double sigma) { qs_.initializeAmplitudeRandom(re_, sigma); }
  void initializeRandom(double sigma = 1e-3) {  // AbstractRunner.hpp:96-101
    machineInitialized_ = true;
    qs_.initializeRandom(re_, sigma);
  }
  Machine& getQs() { return qs_; }
  uint32_t getDim() const { return qs_.getDim(); }

  template <class Callback, class Basis, class Hamiltonian>
  void run(Callback&& callback, Basis&& basis, Hamiltonian&& ham, Mode mode = Mode::ALL) {
    if (!machineInitialized_) initializeRandom();
    if (!opt_) setOptimizer("SGD");
    const size_t dA = qs_.getDimAmp(), D = qs_.getDim();
    NGDExact ngdex(qs_, basis, ham);
    for (int ll = 0; ll <= maxIter_; ll++) {  // :72-121
      logger_ << "Epochs: " << ll << std::endl;
      ngdex.constructExact();
      const double currE = ngdex.eloc();
      const RealVector grad = ngdex.energyGrad();
      const double lambda = std::max(lambdaIni_ * std::pow(lambdaDecay_, ll), lambdaMin_);
      RealVector v = ngdex.solve(grad, lambda);  // both blocks, each with its own factor
      if (mode == Mode::PHASE) std::fill(v.begin(), v.begin() + dA, 0.0);
      if (mode == Mode::AMP) std::fill(v.begin() + dA, v.end(), 0.0);
      const RealVector optV = opt_->getUpdate(v);
      qs_.updateParams(optV);
      double ng = 0.0, nv = 0.0;
      for (size_t i = 0; i < D; ++i) {
        ng += grad[i] * grad[i];
        nv += v[i] * v[i];
      }
      callback(ll, currE, std::sqrt(ng), std::sqrt(nv));
    }
  }

 private:
  std::ostream& logger_;
  Machine qs_;
  std::mt19937_64 re_;
  std::unique_ptr<Optimizer<double>> opt_;
  bool machineInitialized_ = false;
  double lambdaIni_ = 1e-3, lambdaDecay_ = 1.0, lambdaMin_ = 1e-4;
  int maxIter_ = 2000;
};

// Utilities/Utility.hpp:64-88 random configurations for randomize()
template <class RandomEngine>
std::vector<int> randomSigma(int n, RandomEngine& re) {
  std::uniform_int_distribution<> uid(0, 1);
  std::vector<int> s(n);
  for (int i = 0; i < n; i++) s[i] = -2 * uid(re) + 1;
  return s;
}
template <class RandomEngine>
std::vector<int> randomSigma(int n, int nup, RandomEngine& re) {
  std::vector<int> s(n, -1);
  for (int i = nup; i < n; i++) s[i] = 1;
  std::shuffle(s.begin(), s.end(), re);
  return s;
}

}  // namespace yannq_amd
