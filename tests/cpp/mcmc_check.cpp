// pofdMCMC end to end (the pofd_affine_mcmc flow, src/pofd_affine_mcmc.cc):
// init file + spec file -> initial positions -> burn-in -> stored steps on
// the GPU, then the statistics, the text chain, and the last step of walker
// 0 re-evaluated through calcLike.
//
// usage: mcmc_check <initfile> <specfile> <nwalkers> <nsamples> <min_burn>
//                   <chain out> [init_steps fixed_burn] [h5=<file>]
// Prints "npar", "fit", "bonus", "acc <mean acceptance>", one
// "stat <i> <mean> <var> <low> <up>" per parameter, "acor ..." and
// "recheck <chain logL> <direct logL>"; with h5= also writes the run as
// pofdMCMC::writeToHDF5 does.
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/pofd_affine_amd_mcmc.hpp"

int main(int argc, char** argv) {
  std::string h5;
  if (argc > 1 && std::string(argv[argc - 1]).rfind("h5=", 0) == 0) {
    h5 = std::string(argv[argc - 1]).substr(3);
    --argc;
  }
  if (argc != 7 && argc != 9) {
    std::cerr << "usage: mcmc_check init spec nwalkers nsamples min_burn chain "
                 "[init_steps fixed_burn]"
              << std::endl;
    return 2;
  }
  try {
    const unsigned nw = std::stoul(argv[3]), ns = std::stoul(argv[4]), mb = std::stoul(argv[5]);
    const unsigned ninit = (argc == 9) ? std::stoul(argv[7]) : 0;
    const bool fixed = (argc == 9) && std::stoi(argv[8]) != 0;
    pofdMCMC mc(argv[1], argv[2], nw, ns, ninit, 2.0, mb, fixed, 5.0f, 2.0f);
    mc.setSeed(17);
    mc.sample();
    const unsigned np = mc.getNParams();
    std::printf("npar %u\nfit %u\nbonus %u\n", np, mc.getNFitParams(), mc.getNBonusParams());
    std::vector<float> acc;
    mc.getAcceptanceFrac(acc);
    double a = 0;
    for (float v : acc) a += v;
    std::printf("acc %.6f\n", a / acc.size());
    for (unsigned i = 0; i < np; ++i) {
      float mn, var, lo, up;
      mc.getParamStats(i, mn, var, lo, up);
      std::printf("stat %u %.9g %.9g %.9g %.9g %d\n", i, mn, var, lo, up,
                  mc.isParamFixed(i) ? 1 : 0);
    }
    std::vector<float> tau;
    const bool ok = mc.getAcor(tau);
    std::printf("acor %d", ok ? 1 : 0);
    for (float t : tau) std::printf(" %.6g", t);
    std::printf("\n");
    std::vector<float> th;
    std::vector<double> ll;
    mc.getChain(th, ll);
    const size_t last = ll.size() / nw - 1;
    paramSet p(np, th.data() + last * nw * np);
    bool rej = false;
    const double direct = mc.getLogLike(p, rej);
    std::printf("recheck %.17g %.17g\n", ll[last * nw], direct);
    std::printf("maxlike %.17g\n", mc.getMaxLogLike());
    mc.writeToFile(argv[6]);
    if (!h5.empty()) {
#ifdef POFD_AFFINE_HAVE_HDF5
      mc.writeToHDF5(h5);
#else
      std::cerr << "built without HDF5" << std::endl;
      return 1;
#endif
    }
    mc.printStatistics(0.683, std::cout);
  } catch (const affineExcept& e) {
    std::cerr << "affineExcept: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
