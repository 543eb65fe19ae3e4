// Drives the reference-shaped C++ API (include/pofd_affine_amd.hpp) the way
// pofd_affine_getLike does (src/pofd_affine_getLike.cc:16-106): read data
// and beam files, set knots and R ranges, evaluate log likelihoods.
//
// usage: getlike_check fftsize ninterp binned nbins minbeamval nbeamhist
//                      sigma datafile beamfile knotfile thetafile [pdfile]
// knotfile: knot positions (Jy), one per line; thetafile: one parameter
// set per line.  Prints "loglike <value>" per parameter set, then one line
// "pd <minflux> <dflux> <n> <peak log2>" from PDFactory::getPD for the
// first set, "pdlike <value>" from PD::getLogLike of the data given that
// P(D) (written as raw doubles to pdfile when given), and "cache <same>"
// comparing a factory reused for the second set with a fresh one.
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
  if (argc != 12 && argc != 13) {
    std::cerr << "bad usage" << std::endl;
    return 2;
  }
  try {
    const unsigned fftsize = std::stoul(argv[1]), ninterp = std::stoul(argv[2]);
    const bool binned = std::stoi(argv[3]) != 0;
    const unsigned nbins = std::stoul(argv[4]);
    const double minbeamval = std::stod(argv[5]);
    const unsigned nbeamhist = std::stoul(argv[6]);
    const double sigma = std::stod(argv[7]);
    std::vector<double> knots;
    for (auto& r : read_rows(argv[10])) knots.push_back(r[0]);
    auto thetas = read_rows(argv[11]);

    calcLike likeSet(fftsize, ninterp, binned, nbins);
    likeSet.readDataFromFiles({argv[8]}, {argv[9]}, {sigma}, {1.0}, false,
                              false, minbeamval, nbeamhist > 0, nbeamhist,
                              0.0);
    likeSet.setKnotPositions(knots);
    paramSet p0(thetas[0]);
    likeSet.setRRanges(p0);
    for (auto& t : thetas) {
      paramSet p(t);
      bool rej = false;
      double ll = likeSet.getLogLike(p, rej);
      std::printf("loglike %.17g %d\n", ll, rej ? 1 : 0);
    }
    // CFIRB and Poisson priors on the first set, and its bonus parameters
    likeSet.setCFIRBPrior(7000.0, 300.0);
    likeSet.setPoissonPrior(20.0, 2.0);
    {
      bool rej = false;
      double ll = likeSet.getLogLike(p0, rej);
      paramSet pb(p0.getNParams() < knots.size() + 3 ? knots.size() + 3 : p0.getNParams());
      for (unsigned i = 0; i < p0.getNParams(); ++i) pb[i] = p0[i];
      likeSet.fillBonusParams(pb);
      std::printf("prior %.17g %.9g %.9g\n", ll, pb[knots.size() + 1],
                  pb[knots.size() + 2]);
    }
    likeSet.unsetCFIRBPrior();
    likeSet.unsetPoissonPrior();
    // PDFactory::getPD on the same beam and model
    beam bm(argv[9], nbeamhist > 0, nbeamhist, minbeamval);
    numberCountsKnotsSpline model(knots);
    model.setParams(p0);
    dblpair rr = model.getRRange(bm);
    PDFactory pfac(ninterp);
    pfac.initPD(fftsize, 1.05 * rr.first, 1.05 * rr.second + 7.5 * sigma,
                model, bm);
    PD pd;
    pfac.getPD(sigma, pd, true);
    double peak = pd.pd_[0];
    for (unsigned i = 1; i < pd.getDim(); ++i)
      if (pd.pd_[i] > peak) peak = pd.pd_[i];
    std::printf("pd %.17g %.17g %u %.17g\n", pd.minflux, pd.dflux,
                pd.getDim(), peak);
    // PD::getLogLike of the data given that P(D) (binned as the likelihood)
    fitsData data(argv[8], false, false);
    if (binned) data.applyBinning(nbins);
    std::printf("pdlike %.17g\n", pd.getLogLike(data));
    if (argc == 13) {
      std::FILE* f = std::fopen(argv[12], "wb");
      if (!f || std::fwrite(pd.pd_, sizeof(double), pd.getDim(), f) != pd.getDim()) return 5;
      std::fclose(f);
    }
    // the factory's context is kept across initPD calls that change only
    // the knot values: a second set through it equals a fresh factory's
    if (thetas.size() > 1) {
      numberCountsKnotsSpline m2(knots);
      m2.setParams(paramSet(thetas[1]));
      pfac.initPD(fftsize, 1.05 * rr.first, 1.05 * rr.second + 7.5 * sigma, m2, bm);
      PD a, b;
      pfac.getPD(sigma, a, true);
      PDFactory fresh(ninterp);
      fresh.initPD(fftsize, 1.05 * rr.first, 1.05 * rr.second + 7.5 * sigma, m2, bm);
      fresh.getPD(sigma, b, true);
      bool same = a.getDim() == b.getDim() && a.minflux == b.minflux && a.dflux == b.dflux &&
                  std::memcmp(a.pd_, b.pd_, a.getDim() * sizeof(double)) == 0;
      bool differs = std::memcmp(a.pd_, pd.pd_, a.getDim() * sizeof(double)) != 0;
      std::printf("cache %d %d\n", same ? 1 : 0, differs ? 1 : 0);
    }
    // an error path: a NaN knot value must raise affineExcept
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
