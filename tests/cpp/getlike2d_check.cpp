// Drives the two-band C++ API (include/pofd_affine_amd.hpp) the way
// pofd_affine_getLike does for a two-band spec (src/pofd_affine_getLike.cc:
// 108-215): read the map and beam pairs, set the knot, sigma and offset
// positions and the R ranges, evaluate log likelihoods.
//
// usage: getlike2d_check fftsize nedge binned nbins minbeamval nbeamhist
//                        sigma1 sigma2 map1 map2 beam1 beam2 posfile thetafile
// posfile: three lines (knot, sigma-knot and offset-knot positions, Jy);
// thetafile: one parameter set per line.  Prints "loglike <value>" per set,
// then "pd <minflux1> <minflux2> <dflux1> <dflux2> <n1> <n2> <peak>" from
// PDFactoryDouble::getPD for the first set, then the class::method of the
// error a NaN knot value raises.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/pofd_affine_amd.hpp"

static std::vector<std::vector<double>> read_rows(const char* path) {
  std::ifstream in(path);
  std::vector<std::vector<double>> rows;
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream is(line);
    std::vector<double> r;
    double v;
    while (is >> v) r.push_back(v);
    if (!r.empty()) rows.push_back(r);
  }
  return rows;
}

int main(int argc, char** argv) {
  if (argc != 15) {
    std::cerr << "bad usage" << std::endl;
    return 2;
  }
  try {
    const unsigned fftsize = std::stoul(argv[1]), nedge = std::stoul(argv[2]);
    const bool binned = std::stoi(argv[3]) != 0;
    const unsigned nbins = std::stoul(argv[4]);
    const double minbeamval = std::stod(argv[5]);
    const unsigned nbeamhist = std::stoul(argv[6]);
    const double sigma1 = std::stod(argv[7]), sigma2 = std::stod(argv[8]);
    auto pos = read_rows(argv[13]);
    auto thetas = read_rows(argv[14]);
    if (pos.size() != 3) {
      std::cerr << "position file needs 3 lines" << std::endl;
      return 2;
    }

    calcLikeDouble likeSet(fftsize, nedge, true, binned, nbins);
    likeSet.readDataFromFiles({argv[9]}, {argv[10]}, {argv[11]}, {argv[12]},
                              {sigma1}, {sigma2}, {1.0}, false, false,
                              minbeamval, nbeamhist > 0, nbeamhist, 0.0, 0.0);
    likeSet.setPositions(pos[0], pos[1], pos[2]);
    paramSet p0(thetas[0]);
    likeSet.setRRanges(p0);
    for (auto& t : thetas) {
      paramSet p(t);
      bool rej = false;
      double ll = likeSet.getLogLike(p, rej);
      std::printf("loglike %.17g %d\n", ll, rej ? 1 : 0);
    }
    // CFIRB and Poisson priors of both bands on the first set, and its
    // bonus parameters
    likeSet.setCFIRBPrior1(150.0, 20.0);
    likeSet.setCFIRBPrior2(180.0, 25.0);
    likeSet.setPoissonPrior1(2.0, 0.5);
    likeSet.setPoissonPrior2(3.0, 0.7);
    {
      bool rej = false;
      double ll = likeSet.getLogLike(p0, rej);
      const unsigned nmp = pos[0].size() + pos[1].size() + pos[2].size();
      paramSet pb(nmp + 6);
      for (unsigned i = 0; i < nmp + 2 && i < p0.getNParams(); ++i) pb[i] = p0[i];
      likeSet.fillBonusParams(pb);
      std::printf("prior %.17g %.9g %.9g %.9g %.9g\n", ll, pb[nmp + 2], pb[nmp + 3],
                  pb[nmp + 4], pb[nmp + 5]);
    }
    likeSet.unsetCFIRBPrior1();
    likeSet.unsetCFIRBPrior2();
    likeSet.unsetPoissonPrior1();
    likeSet.unsetPoissonPrior2();
    // PDFactoryDouble::getPD on the same beams and model, over the range
    // the likelihood used
    doublebeam bm(argv[11], argv[12], nbeamhist > 0, nbeamhist, minbeamval);
    numberCountsDoubleLogNormal model(pos[0], pos[1], pos[2]);
    model.setParams(p0);
    PDFactoryDouble pfac(nedge);
    pfac.initPD(fftsize, 0.0, 0.5, 0.0, 0.8, model, bm, true);
    PDDouble pd;
    pfac.getPD(sigma1, sigma2, pd, true);
    double peak = pd.pd_[0];
    for (size_t i = 1; i < static_cast<size_t>(pd.n1) * pd.n2; ++i)
      if (pd.pd_[i] > peak) peak = pd.pd_[i];
    std::printf("pd %.17g %.17g %.17g %.17g %u %u %.17g\n", pd.minflux1,
                pd.minflux2, pd.dflux1, pd.dflux2, pd.n1, pd.n2, peak);
    std::vector<double> bad(thetas[0]);
    bad[1] = std::nan("");
    try {
      bool rej = false;
      likeSet.getLogLike(paramSet(bad), rej);
      std::printf("error none\n");
    } catch (const affineExcept& ex) {
      std::printf("error %s::%s\n", ex.getErrClass().c_str(),
                  ex.getErrMethod().c_str());
    }
  } catch (const affineExcept& ex) {
    std::cerr << ex << std::endl;
    return 4;
  }
  return 0;
}
