// Reads the reference's spec and init file fixtures through
// include/pofd_affine_amd_io.hpp and prints every field as "key value"
// lines (lists space separated) for tests/test_io_formats.py to check; also
// draws initial positions and writes a small text chain.
//
// usage: io_check <testdata dir> <output chain path>
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/pofd_affine_amd_io.hpp"

template <typename T>
static void put(const char* k, const std::vector<T>& v) {
  std::cout << k;
  for (const auto& x : v) std::cout << " " << x;
  std::cout << "\n";
}

template <typename T>
static void put(const char* k, const T& v) {
  std::cout << k << " " << v << "\n";
}

template <class F>
static void dump_block(const char* tag, const F& f, unsigned n) {
  for (unsigned i = 0; i < n; ++i) {
    std::cout << tag << " " << i << " fixed " << f.isKnotFixed(i) << " range "
              << f.getKnotRange(i) << " haslow " << f.knotHasLowerLimit(i) << " low "
              << f.getLowerLimit(i) << " hasup " << f.knotHasUpperLimit(i) << " up "
              << f.getUpperLimit(i) << "\n";
  }
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::cerr << "usage: io_check <testdata dir> <chain out>" << std::endl;
    return 2;
  }
  const std::string d = argv[1];
  std::cout.precision(17);
  try {
    specFile s(d + "/specfile_test1D.txt");
    put("s1.datafiles", s.datafiles);
    put("s1.sigmas", s.sigmas);
    put("s1.psffiles", s.psffiles);
    put("s1.like_norm", s.like_norm);
    put("s1.bin_data", s.bin_data);
    put("s1.nbins", s.nbins);
    put("s1.mean_sub", s.mean_sub);
    put("s1.ignore_mask", s.ignore_mask);
    put("s1.fftsize", s.fftsize);
    put("s1.ninterp", s.ninterp);
    put("s1.beam_histogram", s.beam_histogram);
    put("s1.fit_sigma", s.fit_sigma);
    put("s1.has_sigprior", s.has_sigprior);
    put("s1.sigprior_stdev", s.sigprior_stdev);
    put("s1.exp_conf", s.exp_conf);
    put("s1.minbeamval", s.minbeamval);
    put("s1.nbeamhist", s.nbeamhist);
    put("s1.has_cfirbprior", s.has_cfirbprior);
    put("s1.cfirbprior_mean", s.cfirbprior_mean);
    put("s1.cfirbprior_stdev", s.cfirbprior_stdev);
    put("s1.has_poissonprior", s.has_poissonprior);
    put("s1.regularization_alpha", s.regularization_alpha);
    put("s1.has_wisdom_file", s.has_wisdom_file);
    put("s1.wisdom_file", s.wisdom_file);
    put("s1.verbosity", s.verbosity);

    specFileDouble t(d + "/specfile_test2D.txt");
    put("s2.datafiles1", t.datafiles1);
    put("s2.datafiles2", t.datafiles2);
    put("s2.sigmas1", t.sigmas1);
    put("s2.sigmas2", t.sigmas2);
    put("s2.psffiles1", t.psffiles1);
    put("s2.psffiles2", t.psffiles2);
    put("s2.like_norm", t.like_norm);
    put("s2.bin_data", t.bin_data);
    put("s2.nbins", t.nbins);
    put("s2.mean_sub", t.mean_sub);
    put("s2.ignore_mask", t.ignore_mask);
    put("s2.fftsize", t.fftsize);
    put("s2.nedge", t.nedge);
    put("s2.edge_set", t.edge_set);
    put("s2.beam_histogram", t.beam_histogram);
    put("s2.fit_sigma1", t.fit_sigma1);
    put("s2.fit_sigma2", t.fit_sigma2);
    put("s2.has_sigprior1", t.has_sigprior1);
    put("s2.sigprior_stdev1", t.sigprior_stdev1);
    put("s2.has_sigprior2", t.has_sigprior2);
    put("s2.exp_conf1", t.exp_conf1);
    put("s2.exp_conf2", t.exp_conf2);
    put("s2.has_cfirbprior1", t.has_cfirbprior1);
    put("s2.has_cfirbprior2", t.has_cfirbprior2);
    put("s2.cfirbprior_mean2", t.cfirbprior_mean2);
    put("s2.cfirbprior_stdev2", t.cfirbprior_stdev2);
    put("s2.has_poissonprior1", t.has_poissonprior1);
    put("s2.poissonprior_mean1", t.poissonprior_mean1);
    put("s2.poissonprior_stdev1", t.poissonprior_stdev1);
    put("s2.has_poissonprior2", t.has_poissonprior2);
    put("s2.regularization_alpha", t.regularization_alpha);
    put("s2.wisdom_file", t.wisdom_file);
    put("s2.verbosity", t.verbosity);

    initFileKnots f1(d + "/fiducial_1Dmodel.txt", false, false);
    std::vector<double> kp, kv;
    f1.getKnotPos(kp);
    f1.getKnotVals(kv);
    put("i0.nknots", f1.getNKnots());
    put("i0.pos", kp);
    put("i0.val", kv);
    dump_block("i0.k", f1, f1.getNKnots());

    initFileKnots f2(d + "/initfile_test1D.txt", true, true);
    f2.getKnotPos(kp);
    f2.getKnotVals(kv);
    put("i1.nknots", f2.getNKnots());
    put("i1.pos", kp);
    put("i1.val", kv);
    dump_block("i1.k", f2, f2.getNKnots());
    // initial positions: every draw inside the limits and the ranges
    ran rng(12345);
    paramSet p(10);
    int bad = 0;
    double lo_seen = 1e300, hi_seen = -1e300;
    for (int k = 0; k < 20000; ++k) {
      f2.generateRandomKnotValues(rng, p);
      if (!f2.isValid(p)) ++bad;
      for (unsigned i = 0; i < f2.getNKnots(); ++i) {
        const double c = f2.getKnotValue(i), r = f2.getKnotRange(i);
        if (r == 0.0 ? p[i] != static_cast<float>(c)
                     : (p[i] < c - 0.5 * r - 1e-6 || p[i] > c + 0.5 * r + 1e-6))
          ++bad;
      }
      lo_seen = std::min(lo_seen, static_cast<double>(p[0]));
      hi_seen = std::max(hi_seen, static_cast<double>(p[0]));
    }
    put("i1.draw_bad", bad);
    put("i1.draw0_min", lo_seen);
    put("i1.draw0_max", hi_seen);
    paramSet q(10);
    f2.getParams(q);
    q[0] = 2.5f;  // below the lower limit 3
    put("i1.valid_low", f2.isValid(q));
    q[0] = 7.0f;
    put("i1.valid_ok", f2.isValid(q));

    initFileDoubleLogNormal g(d + "/initfile_test2D.txt", true, true);
    put("i2.nknots", g.getNKnots());
    put("i2.nsigmas", g.getNSigmas());
    put("i2.noffsets", g.getNOffsets());
    put("i2.ntot", g.getNTot());
    g.getKnotPos(kp);
    g.getKnotVals(kv);
    put("i2.pos", kp);
    put("i2.val", kv);
    g.getSigmaPos(kp);
    g.getSigmaVals(kv);
    put("i2.sigpos", kp);
    put("i2.sigval", kv);
    g.getOffsetPos(kp);
    g.getOffsetVals(kv);
    put("i2.offpos", kp);
    put("i2.offval", kv);
    dump_block("i2.k", g, g.getNTot());

    // NR3 generator: the first draws for seed 1 and the uniform moments
    ran r1(1);
    std::vector<unsigned long long> first;
    for (int k = 0; k < 3; ++k) first.push_back(r1.int64());
    put("ran.first", first);
    double m1 = 0, m2 = 0;
    const int nr = 1000000;
    for (int k = 0; k < nr; ++k) {
      const double u = r1.doub();
      m1 += u;
      m2 += u * u;
    }
    put("ran.mean", m1 / nr);
    put("ran.var", m2 / nr - (m1 / nr) * (m1 / nr));

    // text chain: 3 stored steps of 2 walkers x 3 parameters
    std::vector<float> th(3 * 2 * 3);
    std::vector<double> ll(3 * 2);
    for (size_t i = 0; i < th.size(); ++i) th[i] = 0.25f * static_cast<float>(i) - 1.0f;
    for (size_t i = 0; i < ll.size(); ++i) ll[i] = -1000.0 - 0.5 * static_cast<double>(i);
    writeChainText(argv[2], th.data(), ll.data(), 3, 2, 3);
  } catch (const affineExcept& e) {
    std::cerr << "affineExcept: " << e.what() << std::endl;
    return 1;
  }
  // malformed inputs are refused with the reference's errors
  const char* bad_lines[] = {"dataset= only_two.fits 0.1\n", "fftsize= -4\n",
                             "no_such_key= 3\n", "sigmaprior= 0\n"};
  int refused = 0;
  for (const char* b : bad_lines) {
    const std::string tmp = std::string(argv[2]) + ".spec";
    FILE* fp = std::fopen(tmp.c_str(), "w");
    std::fputs(b, fp);
    std::fclose(fp);
    try {
      specFile s(tmp);
    } catch (const affineExcept&) {
      ++refused;
    }
    std::remove(tmp.c_str());
  }
  put("bad.refused", refused);
  return 0;
}
