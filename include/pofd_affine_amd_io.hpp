/* Input files and chain output of a fit, with the reference's formats and
   class names, so that its spec files, init files and text chains work
   unchanged: specFile / specFileDouble (include/specFile.h:17-137,
   include/specFileDouble.h), initFileKnots (include/initFileKnots.h,
   src/initFileKnots.cc) and initFileDoubleLogNormal
   (include/initFileDoubleLogNormal.h, src/initFileDoubleLogNormal.cc), the
   NR3 generator the initial positions are drawn with (ran,
   src/ran.cc:8-56), and the text layout of affineChainSet::writeToFile
   (src/affineChainSet.cc:1255-1270, 406-423).  Host code only; HDF5 output
   is not provided (the library is not part of this build).

   Header-only on top of pofd_affine_amd.hpp (paramSet, affineExcept). */
#ifndef POFD_AFFINE_AMD_IO_HPP
#define POFD_AFFINE_AMD_IO_HPP

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "pofd_affine_amd.hpp"

namespace pofd_io {

/* words of a line split at the characters in seps (utility::stringwords /
   stringwords_eq, src/utility.cc:15-81: blanks, tabs and line ends, and
   '=' for spec files) */
inline std::vector<std::string> split_words(const std::string& line, const char* seps) {
  std::vector<std::string> w;
  size_t p = line.find_first_not_of(seps);
  while (p != std::string::npos) {
    const size_t e = line.find_first_of(seps, p);
    w.push_back(line.substr(p, e == std::string::npos ? std::string::npos : e - p));
    if (e == std::string::npos) break;
    p = line.find_first_not_of(seps, e);
  }
  return w;
}

/* utility::string_true (src/utility.cc:88-96): true, t, yes, y in any case */
inline bool string_true(const std::string& s) {
  std::string l(s);
  std::transform(l.begin(), l.end(), l.begin(), [](unsigned char c) { return std::tolower(c); });
  return l == "true" || l == "t" || l == "yes" || l == "y";
}

/* a number the way the reference reads one (operator>> on a stringstream;
   0 when the word is not a number) */
template <typename T>
inline T num(const std::string& w) {
  std::istringstream is(w);
  T v{};
  is >> v;
  return v;
}

/* the lines of a file that carry words: comment lines (a '#' as the first
   character or first word) and blank lines dropped */
inline std::vector<std::pair<std::string, std::vector<std::string>>> read_lines(
    const std::string& path, const char* seps, const char* cls) {
  std::ifstream in(path);
  if (!in) throw affineExcept(cls, "readFile", "Error reading file: " + path);
  std::vector<std::pair<std::string, std::vector<std::string>>> out;
  std::string line;
  while (std::getline(in, line)) {
    if (!line.empty() && line[0] == '#') continue;
    auto w = split_words(line, seps);
    if (w.empty() || w[0][0] == '#') continue;
    out.emplace_back(line, std::move(w));
  }
  return out;
}

inline void need(const std::vector<std::string>& w, size_t n, const std::string& line,
                 const char* cls) {
  if (w.size() < n)
    throw affineExcept(cls, "readFile",
                       w[0] + " line doesn't have right number of entries: " + line);
}

}  // namespace pofd_io

/* ---- 1D fit specification (specFile, src/specFile.cc:20-416).  Lines are
   "identifier= values"; dataset= lines add up, any other identifier keeps
   its last value; an unknown identifier is an error. ---- */
struct specFile {
  std::vector<std::string> datafiles;
  std::vector<double> sigmas;
  std::vector<std::string> psffiles;
  std::vector<double> like_norm;
#ifdef POFD_AFFINE_HAVE_HDF5
  /* specFile::writeToHDF5Handle, src/specFile.cc:406-416 */
  void writeToHDF5Handle(hid_t objid) const {
    pofd_hdf5::need(objid >= 0 && H5Iget_ref(objid) >= 0, "specFile", "writeToHDF5Handle",
                    "Input handle is not valid");
    pofd_hdf5::writeDataStrings(objid, "DataFiles", datafiles);
    pofd_hdf5::writeDataStrings(objid, "BeamFiles", psffiles);
    pofd_hdf5::writeDataDoubles(objid, "InstrumentSigma", sigmas);
    pofd_hdf5::writeDataDoubles(objid, "LikelihoodNormalization", like_norm);
  }
#endif
  bool bin_data;
  unsigned int nbins;
  bool mean_sub;
  bool ignore_mask;
  unsigned int fftsize;
  unsigned int ninterp;
  double minbeamval;
  bool beam_histogram;
  unsigned int nbeamhist;
  bool fit_sigma;
  bool has_sigprior;
  double sigprior_stdev;
  double exp_conf;
  bool has_cfirbprior;
  double cfirbprior_mean, cfirbprior_stdev;
  bool has_poissonprior;
  double poissonprior_mean, poissonprior_stdev;
  double regularization_alpha;
  bool has_wisdom_file;
  std::string wisdom_file;
  unsigned int verbosity;
  bool has_user_seed;
  unsigned long long int seed;

  specFile() { init(); }
  explicit specFile(const std::string& f) { readFile(f); }

  void init() {
    datafiles.clear();
    sigmas.clear();
    psffiles.clear();
    like_norm.clear();
    bin_data = false;
    nbins = 0;
    mean_sub = false;
    ignore_mask = false;
    fftsize = 131072;
    ninterp = 2048;
    minbeamval = 1e-5;
    beam_histogram = true;
    nbeamhist = 120;
    fit_sigma = false;
    has_sigprior = false;
    sigprior_stdev = 0.0;
    exp_conf = 0.0;
    has_cfirbprior = false;
    cfirbprior_mean = cfirbprior_stdev = 0.0;
    has_poissonprior = false;
    poissonprior_mean = poissonprior_stdev = 0.0;
    regularization_alpha = 0.0;
    has_wisdom_file = false;
    wisdom_file.clear();
    verbosity = 0;
    has_user_seed = false;
    seed = 0;
  }

  void readFile(const std::string& flname) {
    using namespace pofd_io;
    static const char* C = "specFile";
    init();
    for (auto& lw : read_lines(flname, " =\t\r\n", C)) {
      const std::string& line = lw.first;
      const std::vector<std::string>& w = lw.second;
      const std::string& key = w[0];
      auto bad = [&](const std::string& what) {
        throw affineExcept(C, "readFile", what + " from line: " + line);
      };
      if (key == "dataset") {
        need(w, 4, line, C);
        const double s = num<double>(w[2]);
        if (s < 0.0) bad("Invalid (negative) sigma");
        datafiles.push_back(w[1]);
        sigmas.push_back(s);
        psffiles.push_back(w[3]);
        like_norm.push_back(w.size() >= 5 ? num<double>(w[4]) : 1.0);
      } else if (key == "bin_data") {
        need(w, 2, line, C);
        const int v = num<int>(w[1]);
        if (v <= 0) bad("Invalid (non-positive) bin_data");
        bin_data = true;
        nbins = static_cast<unsigned>(v);
      } else if (key == "mean_sub") {
        need(w, 2, line, C);
        mean_sub = string_true(w[1]);
      } else if (key == "ignore_mask") {
        need(w, 2, line, C);
        ignore_mask = string_true(w[1]);
      } else if (key == "fftsize" || key == "ninterp" || key == "nbeamhist" ||
                 key == "verbosity") {
        need(w, 2, line, C);
        const int v = num<int>(w[1]);
        if (v <= 0) bad("Invalid (non-positive) " + key);
        if (key == "fftsize") fftsize = static_cast<unsigned>(v);
        else if (key == "ninterp") ninterp = static_cast<unsigned>(v);
        else if (key == "nbeamhist") nbeamhist = static_cast<unsigned>(v);
        else verbosity = static_cast<unsigned>(v);
      } else if (key == "minbeamval") {
        need(w, 2, line, C);
        const double v = num<double>(w[1]);
        if (v < 0.0) bad("Invalid (negative) minbeamval");
        minbeamval = v;
      } else if (key == "beam_histogram") {
        need(w, 2, line, C);
        beam_histogram = string_true(w[1]);
      } else if (key == "fit_sigma") {
        need(w, 2, line, C);
        fit_sigma = string_true(w[1]);
      } else if (key == "sigmaprior") {
        need(w, 2, line, C);
        const double v = num<double>(w[1]);
        if (v <= 0.0) bad("Invalid (non-positive) sigma prior stdev");
        has_sigprior = fit_sigma = true;
        sigprior_stdev = v;
      } else if (key == "exp_conf") {
        need(w, 2, line, C);
        const double v = num<double>(w[1]);
        if (v <= 0.0) bad("Invalid (non-positive) exp_conf");
        exp_conf = v;
      } else if (key == "cfirbprior" || key == "poissonprior") {
        need(w, 3, line, C);
        const double m = num<double>(w[1]), sd = num<double>(w[2]);
        if (m <= 0.0) bad("Invalid (non-positive) " + key + " mean");
        if (sd <= 0.0) bad("Invalid (non-positive) " + key + " stdev");
        if (key == "cfirbprior") {
          has_cfirbprior = true;
          cfirbprior_mean = m;
          cfirbprior_stdev = sd;
        } else {
          has_poissonprior = true;
          poissonprior_mean = m;
          poissonprior_stdev = sd;
        }
      } else if (key == "regularize_alpha") {
        need(w, 2, line, C);
        const double v = num<double>(w[1]);
        if (v < 0.0) bad("Invalid (negative) regularization alpha");
        regularization_alpha = v;
      } else if (key == "wisdom_file") {
        need(w, 2, line, C);
        has_wisdom_file = true;
        wisdom_file = w[1];
      } else if (key == "verbose") {
        need(w, 2, line, C);
        if (string_true(w[1])) verbosity = 1;
      } else if (key == "ultraverbose") {
        need(w, 2, line, C);
        if (string_true(w[1])) verbosity = 2;
      } else if (key == "seed") {
        need(w, 2, line, C);
        const long long v = num<long long>(w[1]);
        if (v <= 0) bad("Invalid (non positive) seed");
        seed = static_cast<unsigned long long>(v);
        has_user_seed = true;
      } else {
        throw affineExcept(C, "readFile", "Couldn't determine line type for: " + line);
      }
    }
  }
};

/* ---- 2D fit specification (specFileDouble).  dataset= takes
   datafile1 datafile2 sigma1 sigma2 psffile1 psffile2 [like_norm]; the
   per-band options carry a 1 or 2 suffix; nedge / edge_set control the
   R edge integrals. ---- */
struct specFileDouble {
  std::vector<std::string> datafiles1, datafiles2;
  std::vector<double> sigmas1, sigmas2;
  std::vector<std::string> psffiles1, psffiles2;
  std::vector<double> like_norm;
  bool bin_data;
  unsigned int nbins;
  bool mean_sub;
  bool ignore_mask;
  unsigned int fftsize;
  bool edge_set;
  unsigned int nedge;
  double minbeamval;
  bool beam_histogram;
  unsigned int nbeamhist;
  bool fit_sigma1, fit_sigma2;
  bool has_sigprior1, has_sigprior2;
  double sigprior_stdev1, sigprior_stdev2;
  double exp_conf1, exp_conf2;
  bool has_cfirbprior1, has_cfirbprior2;
  double cfirbprior_mean1, cfirbprior_stdev1, cfirbprior_mean2, cfirbprior_stdev2;
  bool has_poissonprior1, has_poissonprior2;
  double poissonprior_mean1, poissonprior_stdev1, poissonprior_mean2, poissonprior_stdev2;
  double regularization_alpha;
  bool has_wisdom_file;
  std::string wisdom_file;
  unsigned int verbosity;
  bool has_user_seed;
  unsigned long long int seed;

  specFileDouble() { init(); }
  explicit specFileDouble(const std::string& f) { readFile(f); }

  void init() {
    datafiles1.clear();
    datafiles2.clear();
    sigmas1.clear();
    sigmas2.clear();
    psffiles1.clear();
    psffiles2.clear();
    like_norm.clear();
    bin_data = false;
    nbins = 0;
    mean_sub = false;
    ignore_mask = false;
    fftsize = 4096;
    edge_set = true;
    nedge = 256;
    minbeamval = 1e-6;
    beam_histogram = true;
    nbeamhist = 150;
    fit_sigma1 = fit_sigma2 = false;
    has_sigprior1 = has_sigprior2 = false;
    sigprior_stdev1 = sigprior_stdev2 = 0.0;
    exp_conf1 = exp_conf2 = 0.0;
    has_cfirbprior1 = has_cfirbprior2 = false;
    cfirbprior_mean1 = cfirbprior_stdev1 = cfirbprior_mean2 = cfirbprior_stdev2 = 0.0;
    has_poissonprior1 = has_poissonprior2 = false;
    poissonprior_mean1 = poissonprior_stdev1 = poissonprior_mean2 = poissonprior_stdev2 = 0.0;
    regularization_alpha = 0.0;
    has_wisdom_file = false;
    wisdom_file.clear();
    verbosity = 0;
    has_user_seed = false;
    seed = 0;
  }

  void readFile(const std::string& flname) {
    using namespace pofd_io;
    static const char* C = "specFileDouble";
    init();
    for (auto& lw : read_lines(flname, " =\t\r\n", C)) {
      const std::string& line = lw.first;
      const std::vector<std::string>& w = lw.second;
      const std::string& key = w[0];
      auto bad = [&](const std::string& what) {
        throw affineExcept(C, "readFile", what + " from line: " + line);
      };
      // the band of a suffixed option (0 if none)
      const char last = key.empty() ? '\0' : key.back();
      const int band = (last == '1') ? 1 : (last == '2' ? 2 : 0);
      const std::string base = band ? key.substr(0, key.size() - 1) : key;
      if (key == "dataset") {
        need(w, 7, line, C);
        const double s1 = num<double>(w[3]), s2 = num<double>(w[4]);
        if (s1 < 0.0 || s2 < 0.0) bad("Invalid (negative) sigma");
        datafiles1.push_back(w[1]);
        datafiles2.push_back(w[2]);
        sigmas1.push_back(s1);
        sigmas2.push_back(s2);
        psffiles1.push_back(w[5]);
        psffiles2.push_back(w[6]);
        like_norm.push_back(w.size() >= 8 ? num<double>(w[7]) : 1.0);
      } else if (band && base == "fit_sigma") {
        need(w, 2, line, C);
        (band == 1 ? fit_sigma1 : fit_sigma2) = string_true(w[1]);
      } else if (band && base == "sigmaprior") {
        need(w, 2, line, C);
        const double v = num<double>(w[1]);
        if (v <= 0.0) bad("Invalid (non-positive) sigma prior stdev");
        if (band == 1) {
          has_sigprior1 = fit_sigma1 = true;
          sigprior_stdev1 = v;
        } else {
          has_sigprior2 = fit_sigma2 = true;
          sigprior_stdev2 = v;
        }
      } else if (band && base == "exp_conf") {
        need(w, 2, line, C);
        const double v = num<double>(w[1]);
        if (v <= 0.0) bad("Invalid (non-positive) exp_conf");
        (band == 1 ? exp_conf1 : exp_conf2) = v;
      } else if (band && (base == "cfirbprior" || base == "poissonprior")) {
        need(w, 3, line, C);
        const double m = num<double>(w[1]), sd = num<double>(w[2]);
        if (m <= 0.0) bad("Invalid (non-positive) " + base + " mean");
        if (sd <= 0.0) bad("Invalid (non-positive) " + base + " stdev");
        if (base == "cfirbprior") {
          (band == 1 ? has_cfirbprior1 : has_cfirbprior2) = true;
          (band == 1 ? cfirbprior_mean1 : cfirbprior_mean2) = m;
          (band == 1 ? cfirbprior_stdev1 : cfirbprior_stdev2) = sd;
        } else {
          (band == 1 ? has_poissonprior1 : has_poissonprior2) = true;
          (band == 1 ? poissonprior_mean1 : poissonprior_mean2) = m;
          (band == 1 ? poissonprior_stdev1 : poissonprior_stdev2) = sd;
        }
      } else if (key == "bin_data") {
        need(w, 2, line, C);
        const int v = num<int>(w[1]);
        if (v <= 0) bad("Invalid (non-positive) bin_data");
        bin_data = true;
        nbins = static_cast<unsigned>(v);
      } else if (key == "mean_sub") {
        need(w, 2, line, C);
        mean_sub = string_true(w[1]);
      } else if (key == "ignore_mask") {
        need(w, 2, line, C);
        ignore_mask = string_true(w[1]);
      } else if (key == "edge_set") {
        need(w, 2, line, C);
        edge_set = string_true(w[1]);
      } else if (key == "fftsize" || key == "nedge" || key == "nbeamhist" ||
                 key == "verbosity") {
        need(w, 2, line, C);
        const int v = num<int>(w[1]);
        if (v <= 0) bad("Invalid (non-positive) " + key);
        if (key == "fftsize") fftsize = static_cast<unsigned>(v);
        else if (key == "nedge") nedge = static_cast<unsigned>(v);
        else if (key == "nbeamhist") nbeamhist = static_cast<unsigned>(v);
        else verbosity = static_cast<unsigned>(v);
      } else if (key == "minbeamval") {
        need(w, 2, line, C);
        const double v = num<double>(w[1]);
        if (v < 0.0) bad("Invalid (negative) minbeamval");
        minbeamval = v;
      } else if (key == "beam_histogram") {
        need(w, 2, line, C);
        beam_histogram = string_true(w[1]);
      } else if (key == "regularize_alpha") {
        need(w, 2, line, C);
        const double v = num<double>(w[1]);
        if (v < 0.0) bad("Invalid (negative) regularization alpha");
        regularization_alpha = v;
      } else if (key == "wisdom_file") {
        need(w, 2, line, C);
        has_wisdom_file = true;
        wisdom_file = w[1];
      } else if (key == "verbose") {
        need(w, 2, line, C);
        if (string_true(w[1])) verbosity = 1;
      } else if (key == "ultraverbose") {
        need(w, 2, line, C);
        if (string_true(w[1])) verbosity = 2;
      } else if (key == "seed") {
        need(w, 2, line, C);
        const long long v = num<long long>(w[1]);
        if (v <= 0) bad("Invalid (non positive) seed");
        seed = static_cast<unsigned long long>(v);
        has_user_seed = true;
      } else {
        throw affineExcept(C, "readFile", "Couldn't determine line type for: " + line);
      }
    }
  }
};

/* ---- The generator of Numerical Recipes 3rd ed., section 7.1.3 ("Ran":
   a 64-bit MLCG, a 64-bit xorshift and a multiply-with-carry, combined),
   which the reference uses for its initial positions. ---- */
class ran {
  uint64_t u_, v_, w_;

 public:
  explicit ran(uint64_t seed = 10214ULL) { set_seed(seed); }
  void set_seed(uint64_t j) {
    v_ = 4101842887655102017ULL;
    w_ = 1;
    u_ = j ^ v_;
    int64();
    v_ = u_;
    int64();
    w_ = v_;
    int64();
  }
  uint64_t int64() {
    u_ = u_ * 2862933555777941757ULL + 7046029254386353087ULL;
    v_ ^= v_ >> 17;
    v_ ^= v_ << 31;
    v_ ^= v_ >> 8;
    w_ = 4294957665U * (w_ & 0xffffffffULL) + (w_ >> 32);
    uint64_t x = u_ ^ (u_ << 21);
    x ^= x >> 35;
    x ^= x << 4;
    return (x + v_) ^ w_;
  }
  double doub() { return 5.42101086242752217E-20 * static_cast<double>(int64()); }
  float flt() { return static_cast<float>(doub()); }
};

namespace pofd_io {

/* one parameter block of an init file: position, central value, optional
   range (0: fixed) and optional lower / upper limits per entry
   (initFileKnots::readFile, src/initFileKnots.cc:120-233) */
struct KnotBlock {
  std::vector<double> pos, val, range, low, up;
  std::vector<bool> has_low, has_up;
  bool has_range = false, has_lower = false, has_upper = false;

  void add(const std::vector<std::string>& w, bool read_range, bool read_limits) {
    pos.push_back(num<double>(w[0]));
    val.push_back(num<double>(w[1]));
    const double nan = std::numeric_limits<double>::quiet_NaN();
    if (read_range) {
      has_range = true;
      range.push_back(num<double>(w[2]));
    }
    if (!read_limits) return;
    // limits are ignored for a fixed entry (range 0)
    if (w.size() > 3 && !(has_range && range.back() == 0)) {
      has_lower = true;
      has_low.push_back(true);
      low.push_back(num<double>(w[3]));
      if (w.size() > 4) {
        has_upper = true;
        has_up.push_back(true);
        up.push_back(num<double>(w[4]));
      } else {
        has_up.push_back(false);
        up.push_back(nan);
      }
    } else {
      has_low.push_back(false);
      low.push_back(nan);
      has_up.push_back(false);
      up.push_back(nan);
    }
  }

  size_t size() const { return pos.size(); }
  bool lowlim(size_t i) const { return has_lower && has_low[i]; }
  bool uplim(size_t i) const { return has_upper && has_up[i]; }

  /* checkLimitsDontCross and checkRange (src/initFileKnots.cc:245-300) */
  void check(const char* cls) const {
    for (size_t i = 0; i < size(); ++i) {
      if (lowlim(i) && uplim(i)) {
        if (up[i] < low[i]) {
          std::ostringstream e;
          e << "Lower/Upper limits cross at index: " << i << " Lower limit: " << low[i]
            << " Upper limit: " << up[i];
          throw affineExcept(cls, "checkLimitsDontCross", e.str());
        }
        if (has_range && range[i] > 0.0 && up[i] == low[i]) {
          std::ostringstream e;
          e << "Lower/Upper limits meet at index: " << i << " but range is not zero";
          throw affineExcept(cls, "checkLimitsDontCross", e.str());
        }
      }
      if (has_range && range[i] == 0.0) {
        if (lowlim(i) && val[i] < low[i]) {
          std::ostringstream e;
          e << "At knot " << i << " range is zero but mean value " << val[i]
            << " lies below lower limit " << low[i];
          throw affineExcept(cls, "checkRange", e.str());
        }
        if (uplim(i) && val[i] > up[i]) {
          std::ostringstream e;
          e << "At knot " << i << " range is zero but mean value " << val[i]
            << " lies above upper limit " << up[i];
          throw affineExcept(cls, "checkRange", e.str());
        }
      }
    }
  }

  /* generateRandomKnotValues (src/initFileKnots.cc:361-526): uniform in
     [central - range/2, central + range/2), redrawn (up to 1000 times)
     while outside the limits; uniform between the limits when they are
     closer together than the range; fixed entries keep the central value */
  template <class RNG>
  void draw(RNG& rng, const double* cen, double* out, size_t off, const char* cls) const {
    const char* M = "generateRandomKnotValues";
    for (size_t i = 0; i < size(); ++i) {
      const double c = cen[off + i];
      if (!has_range) {
        if (lowlim(i) && c < low[i]) throw affineExcept(cls, M, "central value below lower limit");
        if (uplim(i) && c > up[i]) throw affineExcept(cls, M, "central value above upper limit");
        out[off + i] = c;
        continue;
      }
      const double r = range[i];
      if (!(r > 0.0)) {
        out[off + i] = c;
        continue;
      }
      if (lowlim(i) && low[i] > c + r)
        throw affineExcept(cls, M, "Lower limit is too far above central value");
      if (uplim(i) && up[i] < c - r)
        throw affineExcept(cls, M, "Upper limit is too far below central value");
      if (lowlim(i) && uplim(i) && up[i] - low[i] < r) {
        out[off + i] = low[i] + (up[i] - low[i]) * rng.flt();
        continue;
      }
      for (unsigned it = 0;; ++it) {
        if (it >= 1000) throw affineExcept(cls, M, "Failed to generate acceptable value");
        const double t = rng.flt() * r + (c - 0.5 * r);
        if ((!lowlim(i) || t >= low[i]) && (!uplim(i) || t <= up[i])) {
          out[off + i] = t;
          break;
        }
      }
    }
  }

  bool valid(const paramSet& p, size_t off) const {
    for (size_t i = 0; i < size(); ++i) {
      const double v = p[static_cast<unsigned>(off + i)];
      if (lowlim(i) && v < low[i]) return false;
      if (uplim(i) && v > up[i]) return false;
    }
    return true;
  }
};

}  // namespace pofd_io

/* ---- 1D init file (initFileKnots): per line "knotpos knotval [range
   [lowlim [uplim]]]"; read_range / read_limits select the optional
   columns (limits imply the range). ---- */
class initFileKnots {
  pofd_io::KnotBlock b_;
  static constexpr const char* C = "initFileKnots";
  void idx_ok(unsigned i, const char* m) const {
    if (i >= b_.size()) throw affineExcept(C, m, "Invalid knot index");
  }

 public:
  initFileKnots() = default;
  initFileKnots(const std::string& f, bool read_range = false, bool read_limits = false) {
    readFile(f, read_range, read_limits);
  }
  void readFile(const std::string& flname, bool read_range = false, bool read_limits = false) {
    if (read_limits) read_range = true;
    b_ = pofd_io::KnotBlock();
    const size_t nreq = read_range ? 3 : 2;
    for (auto& lw : pofd_io::read_lines(flname, " \t\r\n", C))
      if (lw.second.size() >= nreq) b_.add(lw.second, read_range, read_limits);
    if (b_.size() == 0)
      throw affineExcept(C, "readFiles", "No knot positions or values found in " + flname);
    b_.check(C);
  }
  unsigned int getNKnots() const { return static_cast<unsigned>(b_.size()); }
  void getKnotPos(std::vector<double>& v) const { v = b_.pos; }
  void getKnotVals(std::vector<double>& v) const { v = b_.val; }
  double getKnotPos(unsigned i) const {
    idx_ok(i, "getKnotPos");
    return b_.pos[i];
  }
  double getKnotValue(unsigned i) const {
    idx_ok(i, "getKnotValue");
    return b_.val[i];
  }
  double getKnotRange(unsigned i) const {
    idx_ok(i, "getKnotRange");
    return b_.has_range ? b_.range[i] : 0.0;
  }
  bool isKnotFixed(unsigned i) const {
    idx_ok(i, "isKnotFixed");
    return b_.has_range && b_.range[i] == 0.0;
  }
  bool knotHasLowerLimit(unsigned i) const {
    idx_ok(i, "knotHasLowerLimit");
    return b_.lowlim(i);
  }
  double getLowerLimit(unsigned i) const {
    idx_ok(i, "getLowerLimit");
    return b_.lowlim(i) ? b_.low[i] : std::numeric_limits<double>::quiet_NaN();
  }
  bool knotHasUpperLimit(unsigned i) const {
    idx_ok(i, "knotHasUpperLimit");
    return b_.uplim(i);
  }
  double getUpperLimit(unsigned i) const {
    idx_ok(i, "getUpperLimit");
    return b_.uplim(i) ? b_.up[i] : std::numeric_limits<double>::quiet_NaN();
  }
  bool hasRange() const { return b_.has_range; }
  bool hasLowerLimits() const { return b_.has_lower; }
  bool hasUpperLimits() const { return b_.has_upper; }
#ifdef POFD_AFFINE_HAVE_HDF5
  /* initFileKnots::writeToHDF5Handle, src/initFileKnots.cc:627-655 (limits
     a knot does not have are written as NaN) */
  void writeToHDF5Handle(hid_t objid) const {
    using namespace pofd_hdf5;
    need(objid >= 0 && H5Iget_ref(objid) >= 0, C, "writeToHDF5Handle",
         "Input handle is not valid");
    writeAttBool(objid, "HasInitParamRange", b_.has_range);
    writeAttBool(objid, "HasParamLowerLimits", b_.has_lower);
    writeAttBool(objid, "HasParamUpperLimits", b_.has_upper);
    const unsigned n = getNKnots();
    if (n == 0) return;
    if (b_.has_range) writeDataDoubles(objid, "ParamInitRange", b_.range);
    std::vector<bool> has(n);
    std::vector<double> lim(n);
    if (b_.has_lower) {
      for (unsigned i = 0; i < n; ++i) {
        has[i] = knotHasLowerLimit(i);
        lim[i] = getLowerLimit(i);
      }
      writeDataBools(objid, "ParamHasLowerlim", has);
      writeDataDoubles(objid, "ParamLowerlim", lim);
    }
    if (b_.has_upper) {
      for (unsigned i = 0; i < n; ++i) {
        has[i] = knotHasUpperLimit(i);
        lim[i] = getUpperLimit(i);
      }
      writeDataBools(objid, "ParamHasUpperlim", has);
      writeDataDoubles(objid, "ParamUpperlim", lim);
    }
  }
#endif
  /* central values into p[0..nknots) */
  void getParams(paramSet& p) const {
    if (p.getNParams() < b_.size()) throw affineExcept(C, "getParams", "Not enough space in provided paramSet");
    for (size_t i = 0; i < b_.size(); ++i) p[static_cast<unsigned>(i)] = static_cast<float>(b_.val[i]);
  }
  template <class RNG>
  void generateRandomKnotValues(RNG& rng, paramSet& pnew) const {
    paramSet pcen(pnew.getNParams());
    getParams(pcen);
    generateRandomKnotValues(rng, pnew, pcen);
  }
  template <class RNG>
  void generateRandomKnotValues(RNG& rng, paramSet& pnew, const paramSet& pcen) const {
    if (pnew.getNParams() < b_.size() || pcen.getNParams() < b_.size())
      throw affineExcept(C, "generateRandomKnotValues", "Not enough space in provided paramSet");
    std::vector<double> cen(b_.size()), out(b_.size());
    for (size_t i = 0; i < b_.size(); ++i) cen[i] = pcen[static_cast<unsigned>(i)];
    b_.draw(rng, cen.data(), out.data(), 0, C);
    for (size_t i = 0; i < b_.size(); ++i) pnew[static_cast<unsigned>(i)] = static_cast<float>(out[i]);
  }
  bool isValid(const paramSet& p) const {
    if (!(b_.has_lower || b_.has_upper)) return true;
    if (p.getNParams() < b_.size()) throw affineExcept(C, "isValid", "Not enough params in paramSet to test validity");
    return b_.valid(p, 0);
  }
};

/* ---- 2D init file (initFileDoubleLogNormal): a first line
   "nknots nsigmas noffsets", then that many band-1 knot lines, colour
   sigma lines and colour offset lines in order, each as in the 1D file.
   Entries are indexed over all three blocks (0 .. ntot-1). ---- */
class initFileDoubleLogNormal {
  pofd_io::KnotBlock b_;
  unsigned nk_ = 0, ns_ = 0, no_ = 0;
  static constexpr const char* C = "initFileDoubleLogNormal";
  void idx_ok(unsigned i, const char* m) const {
    if (i >= b_.size()) throw affineExcept(C, m, "Invalid index");
  }
  std::vector<double> slice(const std::vector<double>& v, unsigned a, unsigned n) const {
    return std::vector<double>(v.begin() + a, v.begin() + a + n);
  }

 public:
  initFileDoubleLogNormal() = default;
  initFileDoubleLogNormal(const std::string& f, bool read_range = false, bool read_limits = false) {
    readFile(f, read_range, read_limits);
  }
  void readFile(const std::string& flname, bool read_range = false, bool read_limits = false) {
    if (read_limits) read_range = true;
    b_ = pofd_io::KnotBlock();
    std::ifstream in(flname);
    if (!in) throw affineExcept(C, "readFile", "Unable to open file:" + flname);
    in >> nk_ >> ns_ >> no_;
    if (!in || nk_ < 2) throw affineExcept(C, "readFile", "Need at least 2 band 1 knots");
    if (ns_ < 1) throw affineExcept(C, "readFile", "Need at least one sigma color model knot");
    if (no_ < 1) throw affineExcept(C, "readFile", "Need at least one offset color model knot");
    std::string line;
    std::getline(in, line);  // the rest of the count line
    const size_t nreq = read_range ? 3 : 2;
    while (std::getline(in, line)) {
      if (!line.empty() && line[0] == '#') continue;
      auto w = pofd_io::split_words(line, " \t\r\n");
      if (w.empty() || w[0][0] == '#' || w.size() < nreq) continue;
      b_.add(w, read_range, read_limits);
    }
    if (b_.size() != static_cast<size_t>(nk_) + ns_ + no_) {
      std::ostringstream e;
      e << "Expected " << nk_ + ns_ + no_ << " values, got " << b_.size();
      throw affineExcept(C, "readFile", e.str());
    }
    b_.check(C);
  }
  unsigned int getNKnots() const { return nk_; }
  unsigned int getNSigmas() const { return ns_; }
  unsigned int getNOffsets() const { return no_; }
  unsigned int getNTot() const { return nk_ + ns_ + no_; }
  void getKnotPos(std::vector<double>& v) const { v = slice(b_.pos, 0, nk_); }
  void getKnotVals(std::vector<double>& v) const { v = slice(b_.val, 0, nk_); }
  void getSigmaPos(std::vector<double>& v) const { v = slice(b_.pos, nk_, ns_); }
  void getSigmaVals(std::vector<double>& v) const { v = slice(b_.val, nk_, ns_); }
  void getOffsetPos(std::vector<double>& v) const { v = slice(b_.pos, nk_ + ns_, no_); }
  void getOffsetVals(std::vector<double>& v) const { v = slice(b_.val, nk_ + ns_, no_); }
  double getKnotRange(unsigned i) const {
    idx_ok(i, "getKnotRange");
    return b_.has_range ? b_.range[i] : 0.0;
  }
  bool isKnotFixed(unsigned i) const {
    idx_ok(i, "isKnotFixed");
    return b_.has_range && b_.range[i] == 0.0;
  }
  bool knotHasLowerLimit(unsigned i) const {
    idx_ok(i, "knotHasLowerLimit");
    return b_.lowlim(i);
  }
  double getLowerLimit(unsigned i) const {
    idx_ok(i, "getLowerLimit");
    return b_.lowlim(i) ? b_.low[i] : std::numeric_limits<double>::quiet_NaN();
  }
  bool knotHasUpperLimit(unsigned i) const {
    idx_ok(i, "knotHasUpperLimit");
    return b_.uplim(i);
  }
  double getUpperLimit(unsigned i) const {
    idx_ok(i, "getUpperLimit");
    return b_.uplim(i) ? b_.up[i] : std::numeric_limits<double>::quiet_NaN();
  }
  void getParams(paramSet& p) const {
    if (p.getNParams() < b_.size()) throw affineExcept(C, "getParams", "Not enough space in provided paramSet");
    for (size_t i = 0; i < b_.size(); ++i) p[static_cast<unsigned>(i)] = static_cast<float>(b_.val[i]);
  }
  template <class RNG>
  void generateRandomKnotValues(RNG& rng, paramSet& pnew) const {
    paramSet pcen(pnew.getNParams());
    getParams(pcen);
    generateRandomKnotValues(rng, pnew, pcen);
  }
  /* around the central values pcen instead of the file's */
  template <class RNG>
  void generateRandomKnotValues(RNG& rng, paramSet& pnew, const paramSet& pcen) const {
    if (pnew.getNParams() < b_.size() || pcen.getNParams() < b_.size())
      throw affineExcept(C, "generateRandomKnotValues", "Not enough space in provided paramSet");
    std::vector<double> cen(b_.size()), out(b_.size());
    for (size_t i = 0; i < b_.size(); ++i) cen[i] = pcen[static_cast<unsigned>(i)];
    b_.draw(rng, cen.data(), out.data(), 0, C);
    for (size_t i = 0; i < b_.size(); ++i) pnew[static_cast<unsigned>(i)] = static_cast<float>(out[i]);
  }
  bool isValid(const paramSet& p) const {
    if (!(b_.has_lower || b_.has_upper)) return true;
    if (p.getNParams() < b_.size()) throw affineExcept(C, "isValid", "Not enough params in paramSet to test validity");
    return b_.valid(p, 0);
  }
};

/* ---- Text chain output, the layout of affineChainSet::writeToFile: per
   walker, per stored step, the parameters in scientific notation (the
   first in a 13-wide left-justified field, precision 6), the log
   likelihood (15 wide, precision 8) and the walker index.  theta is
   [nsteps][nwalkers][npar] and loglike [nsteps][nwalkers], as
   pofd_ens_get_chain returns them. ---- */
inline void writeChainText(const std::string& path, const float* theta, const double* loglike,
                           unsigned nsteps, unsigned nwalkers, unsigned npar) {
  std::ofstream os(path);
  if (!os) throw affineExcept("affineChainSet", "writeToFile", "Failed to open file");
  os << std::left << std::scientific << std::showpoint;
  for (unsigned w = 0; w < nwalkers; ++w)
    for (unsigned s = 0; s < nsteps; ++s) {
      const float* p = theta + (static_cast<size_t>(s) * nwalkers + w) * npar;
      os << std::setw(13) << std::setprecision(6) << p[0];
      for (unsigned k = 1; k < npar; ++k) os << " " << p[k];
      os << " " << std::setw(15) << std::setprecision(8)
         << loglike[static_cast<size_t>(s) * nwalkers + w] << " " << w << std::endl;
    }
}

#endif  // POFD_AFFINE_AMD_IO_HPP
