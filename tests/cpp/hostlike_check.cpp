// An affineEnsemble subclass with its own host likelihood, as the
// reference's examples are written (examples/multi_gauss_example.cc): a
// correlated five-dimensional Gaussian centred at 0.5, overriding
// getLogLike, fillBonusParams (a sixth, bonus parameter: the sum of the
// five) and areParamsValid (parameter 0 below 3), sampled by the device
// stretch move through setHostTarget.
//
// usage: hostlike_check nwalkers nsteps seed [throw_after]
// Prints "mean <i> <value>", "var <i> <value>", "bonus_maxdiff <v>",
// "max0 <v>", "accept <frac>"; with throw_after > 0 getLogLike throws on
// that call and the program prints "caught <class>::<method>".
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <vector>

#include "../../include/pofd_affine_amd_mcmc.hpp"

namespace {

// inverse covariance of sigmas {1.0, 0.9, 0.8, 0.7, 0.6} with the band
// correlation {0.1, 0.2, 1, 0.2, 0.1} (computed by the caller of this test
// and passed in the same order as the reference example's file)
class multiGauss : public affineEnsemble {
 public:
  multiGauss(unsigned nw, unsigned nsteps, const std::vector<double>& icov, long throw_after)
      : affineEnsemble(nw, 6, nsteps, 0, 2.0, 100, false), icov_(icov),
        throw_after_(throw_after) {}
  void initChains() override {
    is_init = true;
    setParamBonus(5);
    setHostTarget();
    paramSet p(6);
    for (unsigned i = 0; i < 5; ++i) p[i] = 0.5f;
    generateInitialPosition(p);
  }
  void generateInitialPosition(const paramSet& p) override {
    std::vector<paramSet> pos;
    for (unsigned w = 0; w < getNWalkers(); ++w) {
      paramSet q(6);
      for (unsigned i = 0; i < 5; ++i) q[i] = p[i] + 0.3f * static_cast<float>(gauss());
      q[5] = 0.0f;
      pos.push_back(q);
    }
    setInitialPositions(pos);
  }
  double getLogLike(const paramSet& p, bool& rej) override {
    if (throw_after_ > 0 && ++calls_ >= throw_after_)
      throw affineExcept("multiGauss", "getLogLike", "planned failure");
    rej = false;
    double d[5], s = 0.0;
    for (unsigned i = 0; i < 5; ++i) d[i] = static_cast<double>(p[i]) - 0.5;
    for (unsigned i = 0; i < 5; ++i)
      for (unsigned j = 0; j < 5; ++j) s += d[i] * icov_[i * 5 + j] * d[j];
    return -0.5 * s;
  }
  void fillBonusParams(paramSet& p, bool rej) override {
    if (rej) {
      p[5] = std::nanf("");
      return;
    }
    float t = 0.0f;
    for (unsigned i = 0; i < 5; ++i) t += p[i];
    p[5] = t;
  }
  bool areParamsValid(const paramSet& p) const override { return p[0] < 3.0f; }

 private:
  std::vector<double> icov_;
  long throw_after_, calls_ = 0;
};

}  // namespace

int main(int argc, char** argv) {
  if (argc < 4) return 2;
  const unsigned nw = std::atoi(argv[1]), ns = std::atoi(argv[2]);
  const long throw_after = argc > 4 ? std::atol(argv[4]) : 0;
  std::vector<double> icov(25);
  for (double& v : icov)
    if (!(std::cin >> v)) return 3;
  try {
    multiGauss mg(nw, ns, icov, throw_after);
    mg.setSeed(std::strtoull(argv[3], nullptr, 10));
    mg.setExponentMode(true);  // the bonus parameter is not a dimension
    mg.sample();
    std::vector<float> th;
    std::vector<double> ll;
    mg.getChain(th, ll);
    const size_t rows = ll.size();
    double bmax = 0.0, max0 = -1e30;
    for (size_t r = 0; r < rows; ++r) {
      const float* q = th.data() + r * 6;
      float t = 0.0f;
      for (unsigned i = 0; i < 5; ++i) t += q[i];
      bmax = std::fmax(bmax, std::fabs(static_cast<double>(t) - q[5]));
      max0 = std::fmax(max0, q[0]);
    }
    for (unsigned i = 0; i < 6; ++i) {
      float mn, var, lo, up;
      mg.getParamStats(i, mn, var, lo, up);
      std::printf("mean %u %.9g\nvar %u %.9g\n", i, mn, i, var);
    }
    std::vector<float> acc;
    mg.getAcceptanceFrac(acc);
    double a = 0.0;
    for (float v : acc) a += v;
    std::printf("bonus_maxdiff %.9g\nmax0 %.9g\naccept %.6f\nrows %zu\n", bmax, max0,
                a / acc.size(), rows);
  } catch (const affineExcept& ex) {
    std::printf("caught %s::%s\n", ex.getErrClass().c_str(), ex.getErrMethod().c_str());
    return throw_after > 0 ? 0 : 4;
  }
  return throw_after > 0 ? 5 : 0;
}
