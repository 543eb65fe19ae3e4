// The pofd_affine_getLike flow (src/pofd_affine_getLike.cc:16-106) from the
// reference's input files: a spec file and an init file (knot positions
// and values) through include/pofd_affine_amd_io.hpp, the likelihood
// through the C++ API.  Prints "loglike <value> <rejected>".
//
// usage: speclike_check <specfile> <initfile> <sigma multiplier>
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/pofd_affine_amd_io.hpp"

int main(int argc, char** argv) {
  if (argc != 4) {
    std::cerr << "usage: speclike_check <specfile> <initfile> <sigmult>" << std::endl;
    return 2;
  }
  try {
    initFileKnots model_info(argv[2], false, false);
    specFile spec_info(argv[1]);
    const unsigned nknots = model_info.getNKnots();
    paramSet pars(nknots + 1);
    model_info.getParams(pars);
    pars[nknots] = static_cast<float>(std::stod(argv[3]));
    if (spec_info.datafiles.empty())
      throw affineExcept("pofd_affine_getLike", "getLikeSingle", "No datafiles loaded");
    calcLike likeSet(spec_info.fftsize, spec_info.ninterp, spec_info.bin_data, spec_info.nbins);
    if (spec_info.has_cfirbprior)
      likeSet.setCFIRBPrior(spec_info.cfirbprior_mean, spec_info.cfirbprior_stdev);
    if (spec_info.has_poissonprior)
      likeSet.setPoissonPrior(spec_info.poissonprior_mean, spec_info.poissonprior_stdev);
    if (spec_info.fit_sigma && spec_info.has_sigprior)
      likeSet.setSigmaPrior(spec_info.sigprior_stdev);
    if (spec_info.regularization_alpha > 0.0)
      likeSet.setRegularizationAlpha(spec_info.regularization_alpha);
    likeSet.readDataFromFiles(spec_info.datafiles, spec_info.psffiles, spec_info.sigmas,
                              spec_info.like_norm, spec_info.ignore_mask, spec_info.mean_sub,
                              spec_info.minbeamval, spec_info.beam_histogram,
                              spec_info.nbeamhist, spec_info.exp_conf);
    std::vector<double> kp;
    model_info.getKnotPos(kp);
    likeSet.setKnotPositions(kp);
    likeSet.setRRanges(pars);
    bool rejected = false;
    const double ll = likeSet.getLogLike(pars, rejected);
    std::printf("loglike %.17g %d\n", ll, rejected ? 1 : 0);
  } catch (const affineExcept& e) {
    std::cerr << "affineExcept: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
