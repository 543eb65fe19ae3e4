/*
 * pofd_affine_amd.hpp -- the reference's C++ likelihood API over the C ABI.
 *
 * Header-only layer giving the classes the reference's MPI slaves and CLIs
 * call -- affineExcept, paramSet, beam, fitsData, numberCountsKnotsSpline,
 * PD, PDFactory and calcLike -- with the reference's names, signatures and
 * error behaviour, backed by libpofd_affine_amd.so on a gfx950 GPU.
 * Reference interfaces: include/affineExcept.h, include/paramSet.h:17-71,
 * include/beam.h:28-175, include/fitsData.h:16-96,
 * include/numberCountsKnotsSpline.h:21-70, include/PD.h:24-96,
 * include/PDFactory.h:26-153, include/calcLike.h:193-287.
 *
 * Differences a caller can observe:
 *  - calcLike::getLogLikeBatch evaluates many walkers in one launch;
 *  - the CFIRB and Poisson priors and the bonus parameters
 *    (fillBonusParams) throw affineExcept("...", "...",
 *    "not supported by the device path");
 *  - PD::getLogLike is not provided: log likelihoods come from
 *    calcLike::getLogLike, which runs the whole chain on the device.
 */
#ifndef POFD_AFFINE_AMD_HPP
#define POFD_AFFINE_AMD_HPP

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <ostream>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "pofd_affine_amd.h"

typedef std::pair<double, double> dblpair;

/* include/affineExcept.h */
class affineExcept : public std::exception {
 private:
  std::string errclass, errmethod, errstr;
  mutable std::string whatmsg;

 public:
  explicit affineExcept(const std::string& s) : errstr(s) {}
  affineExcept(const std::string& c, const std::string& m, const std::string& s)
      : errclass(c), errmethod(m), errstr(s) {}
  ~affineExcept() throw() {}
  std::string getErrClass() const { return errclass; }
  std::string getErrMethod() const { return errmethod; }
  std::string getErrStr() const { return errstr; }
  const char* what() const throw() {
    whatmsg = "Error in " + errclass + "::" + errmethod + ": " + errstr;
    return whatmsg.c_str();
  }
};

inline std::ostream& operator<<(std::ostream& os, const affineExcept& ex) {
  return os << ex.what();
}

namespace pofd_amd_detail {
inline int default_device() { return 0; }
inline double logfac() { return log2(10.0); }  /* global_settings.h:69 */
}  // namespace pofd_amd_detail

/* include/paramSet.h: float parameter vector */
class paramSet final {
 private:
  std::vector<float> v;

 public:
  paramSet() {}
  explicit paramSet(unsigned int n) : v(n, 0.0f) {}
  paramSet(const std::vector<float>& x) : v(x) {}
  paramSet(const std::vector<double>& x) : v(x.begin(), x.end()) {}
  paramSet(unsigned int n, const float* const p) : v(p, p + n) {}
  paramSet(unsigned int n, const double* const p) : v(p, p + n) {}
  void setNParams(unsigned int n) { v.assign(n, 0.0f); }
  void clear() { v.clear(); }
  unsigned int getNParams() const noexcept { return static_cast<unsigned>(v.size()); }
  const float& operator[](unsigned int i) const noexcept { return v[i]; }
  float& operator[](unsigned int i) noexcept { return v[i]; }
  void setParamValue(unsigned int i, float val) noexcept { v[i] = val; }
  const float* data() const noexcept { return v.data(); }
};

/* include/beam.h: the beam after readFile / makeHistogram (host, bit-exact) */
class beam {
 private:
  std::shared_ptr<pofd_beam> h;
  double info_[16];
  void refresh() {
    if (h) pofd_beam_info(h.get(), info_);
  }
  void adopt(pofd_beam* b) {
    h.reset(b, pofd_beam_destroy);
    refresh();
  }

 public:
  beam() { std::memset(info_, 0, sizeof(info_)); }
  beam(const std::string& filename, bool histogram = false,
       unsigned int NBINS = 120, double MINVAL = 1e-5) {
    std::memset(info_, 0, sizeof(info_));
    readFile(filename, MINVAL);
    if (histogram) makeHistogram(NBINS);
  }
  /* beam::readFile, src/beam.cc:295-458 */
  void readFile(const std::string& filename, double MINVAL = 1e-5) {
    uint64_t n = 0;
    double pixsize = 0.0;
    if (pofd_fits_read_beam(filename.c_str(), nullptr, 0, &n, &pixsize))
      throw affineExcept("beam", "readFile", pofd_fits_last_error());
    std::vector<double> pix(n);
    if (pofd_fits_read_beam(filename.c_str(), pix.data(), n, &n, &pixsize))
      throw affineExcept("beam", "readFile", pofd_fits_last_error());
    setPixels(pix.data(), n, pixsize, MINVAL);
  }
  /* the same from an in-memory image (FITS storage order) */
  void setPixels(const double* pix, uint64_t n, double pixsize,
                 double MINVAL = 1e-5) {
    pofd_beam* b = nullptr;
    if (pofd_beam_create(&b, pix, n, pixsize, MINVAL))
      throw affineExcept("beam", "readFile", "Beam file has only zero values");
    adopt(b);
  }
  /* beam::makeHistogram, src/beam.cc:467-556 */
  void makeHistogram(unsigned int NBINS = 120) {
    if (!h) throw affineExcept("beam", "makeHistogram", "No beam read");
    if (pofd_beam_make_histogram(h.get(), NBINS))
      throw affineExcept("beam", "makeHistogram", "Invalid (non-positive) nbins");
    refresh();
  }
  bool hasData() const { return hasPos() || hasNeg(); }
  bool hasPos() const { return info_[2] != 0.0; }
  bool hasNeg() const { return info_[3] != 0.0; }
  unsigned int getNPos() const { return static_cast<unsigned>(info_[0]); }
  unsigned int getNNeg() const { return static_cast<unsigned>(info_[1]); }
  double getPixSize() const { return info_[4]; }
  double getMinval() const { return info_[5]; }
  bool isPosHist() const { return info_[6] != 0.0; }
  /* include/beam.h:97 reports the positive flag here; kept as written */
  bool isNegHist() const { return info_[6] != 0.0; }
  unsigned int getNHistPos() const { return static_cast<unsigned>(info_[8]); }
  unsigned int getNHistNeg() const { return static_cast<unsigned>(info_[9]); }
  double getEffectiveArea() const { return info_[10]; }
  double getEffectiveAreaPos() const { return info_[11]; }
  double getEffectiveAreaNeg() const { return info_[12]; }
  double getEffectiveAreaPix() const { return info_[13]; }
  double getEffectiveAreaSq() const { return info_[14]; }
  std::vector<double> getArray(int which) const {
    uint32_t n = 0;
    if (!h) return {};
    pofd_beam_get(h.get(), which, nullptr, &n);
    std::vector<double> out(n);
    if (n) pofd_beam_get(h.get(), which, out.data(), &n);
    return out;
  }
  dblpair getMinMaxPos() const {
    if (!hasPos()) return dblpair(NAN, NAN);
    std::vector<double> p = getArray(0);
    return dblpair(p.front(), p.back());
  }
  dblpair getMinMaxNeg() const {
    if (!hasNeg()) return dblpair(NAN, NAN);
    std::vector<double> p = getArray(1);
    return dblpair(p.front(), p.back());
  }
  const pofd_beam* handle() const { return h.get(); }
};

/* include/fitsData.h: one map, optionally binned (host, bit-exact) */
class fitsData {
 private:
  std::vector<double> data;
  bool is_binned = false;
  unsigned int nbins = 0;
  double bincent0 = 0.0, bindelta = 0.0;
  std::vector<uint32_t> binval;
  std::string file;

 public:
  fitsData() {}
  fitsData(const std::string& f, bool ignore_mask = false, bool meansub = false) {
    readData(f, ignore_mask, meansub);
  }
  /* fitsData::readData, src/fitsData.cc:56-315 */
  void readData(const std::string& f, bool ignore_mask = false,
                bool domeansub = false) {
    uint64_t n = 0;
    if (pofd_fits_read_map(f.c_str(), ignore_mask ? 1 : 0, nullptr, 0, &n))
      throw affineExcept("fitsData", "readData", pofd_fits_last_error());
    data.assign(n, 0.0);
    if (pofd_fits_read_map(f.c_str(), ignore_mask ? 1 : 0, data.data(), n, &n))
      throw affineExcept("fitsData", "readData", pofd_fits_last_error());
    file = f;
    removeBinning();
    if (domeansub) meanSubtract();
  }
  void setData(const double* d, uint64_t n) {
    data.assign(d, d + n);
    removeBinning();
  }
  bool isBinned() const { return is_binned; }
  /* fitsData::applyBinning, src/fitsData.cc:406-437 */
  void applyBinning(unsigned int NBINS) {
    if (data.empty()) return;
    if (NBINS == 0) throw affineExcept("fitsData", "applyBinning", "Trying to bin with no bins");
    if (is_binned && NBINS == nbins) return;
    binval.assign(NBINS, 0u);
    if (pofd_data_bin(data.data(), data.size(), NBINS, binval.data(), &bincent0, &bindelta))
      throw affineExcept("fitsData", "applyBinning", "binning failed");
    nbins = NBINS;
    is_binned = true;
  }
  void removeBinning() {
    is_binned = false;
    nbins = 0;
    binval.clear();
  }
  unsigned int getNBins() const { return nbins; }
  double getBinCent0() const { return bincent0; }
  double getBinDelta() const { return bindelta; }
  bool hasData() const { return !data.empty(); }
  unsigned int getN() const { return static_cast<unsigned>(data.size()); }
  const double* getData() const { return data.data(); }
  const unsigned int* getBinnedData() const {
    return reinterpret_cast<const unsigned int*>(binval.data());
  }
  /* fitsData::meanSubtract, src/fitsData.cc:393-401 */
  double meanSubtract() {
    double m = 0.0;
    if (!data.empty()) pofd_data_mean_subtract(data.data(), data.size(), &m);
    if (is_binned) bincent0 -= m;
    return m;
  }
  dblpair getMinMax() const {
    if (data.empty()) return dblpair(NAN, NAN);
    if (is_binned)
      return dblpair(bincent0, bincent0 + static_cast<double>(nbins - 1) * bindelta);
    double mn = data[0], mx = data[0];
    for (double c : data) {
      if (c < mn) mn = c;
      if (c > mx) mx = c;
    }
    return dblpair(mn, mx);
  }
  double getMin() const { return getMinMax().first; }
  double getMax() const { return getMinMax().second; }
  std::string getFile() const { return file; }
};

/* include/numberCountsKnotsSpline.h: knot positions + float knot values */
class numberCountsKnotsSpline {
 private:
  std::vector<double> knots;
  std::vector<float> values;
  bool loaded = false;

 public:
  numberCountsKnotsSpline() {}
  explicit numberCountsKnotsSpline(const std::vector<double>& k) { setKnotPositions(k); }
  numberCountsKnotsSpline(unsigned int n, const double* const k) { setKnotPositions(n, k); }
  /* numberCountsKnots::setKnotPositions, src/numberCountsKnots.cc:198-238 */
  void setKnotPositions(unsigned int n, const double* const S) {
    for (unsigned i = 0; i < n; ++i)
      if (S[i] <= 0.0)
        throw affineExcept("numberCountsKnots", "setKnots",
                           "Negative knot positions not allowed");
    for (unsigned i = 1; i < n; ++i)
      if (S[i] <= S[i - 1])
        throw affineExcept("numberCountsKnots", "setKnotPositions",
                           "Knot positions not monotonically increasing");
    knots.assign(S, S + n);
    loaded = false;
  }
  void setKnotPositions(const std::vector<double>& S) {
    setKnotPositions(static_cast<unsigned>(S.size()), S.data());
  }
  void setKnotPositions(const std::vector<float>& S) {
    std::vector<double> d(S.begin(), S.end());
    setKnotPositions(d);
  }
  /* numberCountsKnotsSpline::setParams, src/numberCountsKnotsSpline.cc:265-274 */
  void setParams(const paramSet& F) {
    if (knots.size() > F.getNParams())
      throw affineExcept("numberCountsKnotsSpline", "setKnots",
                         "Number of knot values different than expected");
    values.assign(F.data(), F.data() + knots.size());
    loaded = true;
  }
  /* numberCountsKnots::isValid, src/numberCountsKnots.cc:273-284 */
  bool isValid() const noexcept {
    if (knots.empty() || !loaded) return false;
    for (float v : values)
      if (std::isnan(v)) return false;
    return true;
  }
  unsigned int getNKnots() const noexcept { return static_cast<unsigned>(knots.size()); }
  double getKnotPos(unsigned int i) const noexcept { return knots[i]; }
  double getMinFlux() const noexcept { return knots.empty() ? NAN : knots.front(); }
  double getMaxFlux() const noexcept { return knots.empty() ? NAN : knots.back(); }
  const std::vector<double>& positions() const { return knots; }
  const std::vector<float>& knotValues() const { return values; }
  /* numberCountsKnots::getRRangeInternal, src/numberCountsKnots.cc:345-365 */
  dblpair getRRange(const beam& bm) const {
    if (!isValid()) throw affineExcept("numberCountsKnots", "getRRange", "Model is not valid");
    if (!bm.hasData()) throw affineExcept("numberCountsKnots", "getRRange", "Beam is empty");
    const double maxknot = knots.back();
    if (bm.hasPos())
      return dblpair(bm.hasNeg() ? -maxknot * bm.getMinMaxNeg().second : 0.0,
                     maxknot * bm.getMinMaxPos().second);
    return dblpair(-maxknot * bm.getMinMaxNeg().second, 0.0);
  }
  /* numberCountsKnotsSpline::differenceRegularize, src/numberCountsKnotsSpline.cc:540-559 */
  double differenceRegularize(double alpha) const {
    if (!loaded) return NAN;
    const size_t n = knots.size();
    if (n < 2) return 0.0;
    const double lf = pofd_amd_detail::logfac();
    double pen = 0.0, mean_deriv = 0.0;
    for (size_t i = 1; i < n; ++i)
      mean_deriv += (lf * static_cast<double>(values[i]) - lf * static_cast<double>(values[i - 1])) /
                    (log2(knots[i]) - log2(knots[i - 1]));
    mean_deriv /= static_cast<double>(n - 1);
    for (size_t i = 1; i < n; ++i) {
      const double d = (lf * static_cast<double>(values[i]) - lf * static_cast<double>(values[i - 1])) /
                           (log2(knots[i]) - log2(knots[i - 1])) - mean_deriv;
      pen -= d * d;
    }
    return alpha * pen;
  }
};

/* include/PD.h: a P(D) on a regular flux grid */
class PD {
 private:
  std::vector<double> store;

 public:
  bool logflat = false;
  double minflux = 0.0;
  double dflux = 1.0;
  double* pd_ = nullptr;
  explicit PD(unsigned int N = 0, double MINFLUX = 0.0, double DFLUX = 1.0)
      : store(N, 0.0), minflux(MINFLUX), dflux(DFLUX) {
    pd_ = store.empty() ? nullptr : store.data();
  }
  void resize(unsigned int N) {
    store.resize(N);
    pd_ = store.empty() ? nullptr : store.data();
  }
  unsigned int getDim() const { return static_cast<unsigned>(store.size()); }
  bool isLog() const { return logflat; }
  double operator[](unsigned int i) const { return pd_[i]; }
  double getFluxVal(unsigned int i) const { return minflux + static_cast<double>(i) * dflux; }
  double getPDVal(unsigned int i) const { return pd_[i]; }
  /* PD::getPDVal(double), src/PD.cc:443-455 */
  double getPDVal(double x) const {
    const unsigned n = getDim();
    if (!pd_) return NAN;
    int idx = static_cast<int>((x - minflux) / dflux);
    double maxfluxlm = minflux + static_cast<double>(n - 2) * dflux;
    if (x < minflux) return pd_[0];
    if (x > maxfluxlm) return pd_[n - 1];
    double p0 = pd_[idx];
    double dx = x - (minflux + static_cast<double>(idx) * dflux);
    return p0 + dx / dflux * (pd_[idx + 1] - p0);
  }
  void fill(unsigned int N, double MIN, double DF, const double* const DATA, bool LOG = true) {
    resize(N);
    minflux = MIN;
    dflux = DF;
    logflat = LOG;
    if (N) std::memcpy(pd_, DATA, N * sizeof(double));
  }
};

namespace pofd_amd_detail {
inline void throw_status(pofd_ctx* c, int st, const char* cls, const char* meth) {
  const std::string msg = c ? pofd_last_error(c) : "";
  switch (st) {
    case POFD_E_SPLIT:
      throw affineExcept("PDFactory", "findSplitPoint",
                         "Unable to find split point" + (msg.empty() ? std::string() : ": " + msg));
    case POFD_E_INVALID_MODEL:
      throw affineExcept("PDFactory", "initPD", "Invalid model");
    case POFD_E_MEAN:
      throw affineExcept("PDFactory", "unwrapandNormalizePD", "Un-shift amounts not finite");
    case POFD_E_NOGPU:
      throw affineExcept(cls, meth, "no gfx950 GPU available");
    default:
      throw affineExcept(cls, meth, msg.empty() ? "device path failed" : msg);
  }
}
struct CtxDeleter {
  void operator()(pofd_ctx* c) const { pofd_destroy(c); }
};
typedef std::unique_ptr<pofd_ctx, CtxDeleter> CtxPtr;
inline CtxPtr make_ctx(int device, const char* cls, const char* meth) {
  pofd_ctx* c = nullptr;
  int rc = pofd_create(&c, device);
  if (rc) throw_status(nullptr, rc, cls, meth);
  return CtxPtr(c);
}
}  // namespace pofd_amd_detail

/* include/PDFactory.h: initPD + getPD for one model, one noise level.
   Each getPD runs the fused device kernel for a single walker. */
class PDFactory {
 private:
  unsigned int ninterp;
  unsigned int n = 0;
  double minflux = 0.0, maxflux = 0.0;
  numberCountsKnotsSpline model;
  beam bm;
  bool initialized = false;
  int device = pofd_amd_detail::default_device();

 public:
  explicit PDFactory(unsigned int NINTERP = 2048) : ninterp(NINTERP) {}
  unsigned int getNInterp() const { return ninterp; }
  void setNInterp(unsigned int v) { ninterp = v; }
  void setDevice(int d) { device = d; }
  /* PDFactory::initPD, src/PDFactory.cc:1031-1141 (checks as written) */
  bool initPD(unsigned int N, double MINFLUX, double MAXFLUX,
              const numberCountsKnotsSpline& m, const beam& b) {
    if (N == 0) throw affineExcept("PDFactory", "initPD", "Invalid (zero) n");
    if (N == 1) throw affineExcept("PDFactory", "initPD", "Invalid (==0) n");
    if (MAXFLUX <= 0.0) throw affineExcept("PDFactory", "initPD", "Invalid (non-positive) maxflux");
    if (MINFLUX == MAXFLUX) throw affineExcept("PDFactory", "initPD", "Invalid (0) range between min/max flux");
    if (!m.isValid()) throw affineExcept("PDFactory", "initPD", "Invalid model");
    if (!b.hasData()) throw affineExcept("PDFactory", "initPD", "Empty beam");
    n = N;
    minflux = MINFLUX;
    maxflux = MAXFLUX;
    model = m;
    bm = b;
    initialized = true;
    return true;
  }
  /* PDFactory::getPD, src/PDFactory.cc:1154-1243 */
  void getPD(double sigma, PD& pd, bool setLog = true) {
    if (!initialized) throw affineExcept("PDFactory", "getPD", "Must call initPD first");
    using namespace pofd_amd_detail;
    CtxPtr c = make_ctx(device, "PDFactory", "getPD");
    const std::vector<double>& k = model.positions();
    int rc = pofd1d_configure(c.get(), n, ninterp, k.data(), static_cast<uint32_t>(k.size()), 0.0);
    int b = 0, f = 0;
    if (!rc) rc = pofd1d_add_beamset(c.get(), bm.handle(), &b);
    const uint32_t one = 1;
    if (!rc) rc = pofd1d_add_field(c.get(), b, nullptr, 1, &one, 1, 0.0, 1.0, sigma, 1.0, &f);
    if (!rc) rc = pofd1d_prepare(c.get());
    if (!rc) rc = pofd1d_set_rrange(c.get(), b, minflux, maxflux);
    if (rc) throw_status(c.get(), rc, "PDFactory", "getPD");
    std::vector<float> th(model.knotValues());
    th.push_back(1.0f);
    double ll = 0.0;
    int32_t st = 0;
    rc = pofd1d_loglike_batch(c.get(), th.data(), 1, static_cast<uint32_t>(th.size()), &ll, &st);
    if (rc) throw_status(c.get(), rc, "PDFactory", "getPD");
    if (st) throw_status(c.get(), st, "PDFactory", "getPD");
    pd.resize(n);
    double mf = 0.0, df = 0.0;
    rc = pofd1d_get_pd(c.get(), 0, b, f, setLog ? 1 : 0, pd.pd_, &mf, &df);
    if (rc) throw_status(c.get(), rc, "PDFactory", "getPD");
    pd.minflux = mf;
    pd.dflux = df;
    pd.logflat = setLog;
  }
};

/* include/calcLike.h: data sets grouped by beam, priors, the likelihood */
class calcLike {
 private:
  unsigned int fftsize, ninterp;
  bool bin_data;
  unsigned int nbins;
  mutable numberCountsKnotsSpline model;
  std::vector<fitsData> datasets;
  std::vector<std::pair<beam, std::vector<size_t>>> sets;  // beam, dataset idx
  std::vector<double> sigmas, like_norms;
  double exp_conf = 0.0;
  bool has_sigma_prior = false;
  double sigma_prior_width = 0.0;
  double regularization_alpha = 0.0;
  int device = pofd_amd_detail::default_device();
  mutable pofd_amd_detail::CtxPtr ctx;
  mutable bool prepared = false;

  void build() const {
    using namespace pofd_amd_detail;
    ctx = make_ctx(device, "calcLike", "setRRanges");
    const std::vector<double>& k = model.positions();
    int rc = pofd1d_configure(ctx.get(), fftsize, ninterp, k.data(),
                              static_cast<uint32_t>(k.size()), exp_conf);
    for (size_t s = 0; s < sets.size() && !rc; ++s) {
      int b = 0;
      rc = pofd1d_add_beamset(ctx.get(), sets[s].first.handle(), &b);
      for (size_t i : sets[s].second) {
        if (rc) break;
        const fitsData& d = datasets[i];
        int f = 0;
        if (d.isBinned())
          rc = pofd1d_add_field(ctx.get(), b, nullptr, d.getN(),
                                reinterpret_cast<const uint32_t*>(d.getBinnedData()),
                                d.getNBins(), d.getBinCent0(), d.getBinDelta(),
                                sigmas[i], like_norms[i], &f);
        else
          rc = pofd1d_add_field(ctx.get(), b, d.getData(), d.getN(), nullptr, 0,
                                0.0, 0.0, sigmas[i], like_norms[i], &f);
      }
    }
    if (!rc) rc = pofd1d_prepare(ctx.get());
    if (rc) throw_status(ctx.get(), rc, "calcLike", "setRRanges");
    prepared = true;
  }

 public:
  explicit calcLike(unsigned int FFTSIZE = 262144, unsigned int NINTERP = 1024,
                    bool BINNED = false, unsigned int NBINS = 1000)
      : fftsize(FFTSIZE), ninterp(NINTERP), bin_data(BINNED), nbins(NBINS) {}
  calcLike(const calcLike&) = delete;
  calcLike& operator=(const calcLike&) = delete;

  void setDevice(int d) { device = d; prepared = false; }
  void setFFTSize(unsigned int v) noexcept { fftsize = v; prepared = false; }
  unsigned int getFFTSize() const noexcept { return fftsize; }
  void setNInterp(unsigned int v) { ninterp = v; prepared = false; }
  unsigned int getNInterp() const noexcept { return ninterp; }
  void setBinData() {
    bin_data = true;
    for (auto& d : datasets) d.applyBinning(nbins);
    prepared = false;
  }
  void unSetBinData() {
    bin_data = false;
    for (auto& d : datasets) d.removeBinning();
    prepared = false;
  }
  void setNBins(unsigned int v) {
    nbins = v;
    if (bin_data)
      for (auto& d : datasets) d.applyBinning(nbins);
    prepared = false;
  }

  /* calcLike::readDataFromFiles, src/calcLike.cc:688-758: datasets are
     grouped by beam file; groups are ordered by beam file name (std::map) */
  void readDataFromFiles(const std::vector<std::string>& datafiles,
                         const std::vector<std::string>& beamfiles,
                         const std::vector<double>& sigs,
                         const std::vector<double>& lnorms,
                         bool IGNOREMASK = false, bool MEANSUB = false,
                         double MINBEAMVAL = 1e-5, bool HISTOGRAMBEAMS = false,
                         unsigned int NBEAMHIST = 120, double EXPCONF = 0.0) {
    const size_t ndat = datafiles.size();
    if (ndat == 0) throw affineExcept("calcLike", "readDataFromFiles", "No datafiles");
    if (ndat != beamfiles.size())
      throw affineExcept("calcLike", "readDataFromFiles", "datafiles and beamfiles not same length");
    if (ndat != sigs.size())
      throw affineExcept("calcLike", "readDataFromFiles", "datafiles and sigma not same length");
    if (ndat != lnorms.size())
      throw affineExcept("calcLike", "readDataFromFiles", "datafiles and like_norm not same length");
    std::map<std::string, std::vector<size_t>> groups;
    for (size_t i = 0; i < ndat; ++i) groups[beamfiles[i]].push_back(i);
    datasets.clear();
    sets.clear();
    sigmas.clear();
    like_norms.clear();
    for (auto& g : groups) {
      beam bm(g.first, HISTOGRAMBEAMS, NBEAMHIST, MINBEAMVAL);
      std::vector<size_t> idx;
      for (size_t i : g.second) {
        datasets.emplace_back();
        fitsData& d = datasets.back();
        d.readData(datafiles[i], IGNOREMASK, MEANSUB);
        if (d.getN() == 0)
          throw affineExcept("calcLikeSingle", "readDataFromFile", "No unmasked pixels in data");
        if (bin_data) d.applyBinning(nbins);
        sigmas.push_back(sigs[i]);
        like_norms.push_back(lnorms[i]);
        idx.push_back(datasets.size() - 1);
      }
      sets.emplace_back(bm, idx);
    }
    exp_conf = EXPCONF;
    prepared = false;
  }

  unsigned int getNBeamSets() const noexcept { return static_cast<unsigned>(sets.size()); }
  unsigned int getNKnots() const noexcept { return model.getNKnots(); }
  void setKnotPositions(std::vector<double>& k) { model.setKnotPositions(k); prepared = false; }

  /* calcLike::setRRanges, src/calcLike.cc:827-831 */
  void setRRanges(const paramSet& p) {
    model.setParams(p);
    for (auto& s : sets) (void)model.getRRange(s.first);  // validity checks
    build();
  }
  void setSigmaPrior(double val) noexcept { has_sigma_prior = true; sigma_prior_width = val; }
  void unsetSigmaPrior() noexcept { has_sigma_prior = false; }
  void setCFIRBPrior(double, double) {
    throw affineExcept("calcLike", "setCFIRBPrior", "not supported by the device path");
  }
  void setPoissonPrior(double, double) {
    throw affineExcept("calcLike", "setPoissonPrior", "not supported by the device path");
  }
  void setRegularizationAlpha(double alpha) {
    if (alpha < 0.0)
      throw affineExcept("calcLike", "setRegularizationAlpha", "Invalid (negative) regularization alpha");
    regularization_alpha = alpha;
  }

  /* calcLike::getLogLike, src/calcLike.cc:884-955 */
  double getLogLike(const paramSet& p, bool& pars_invalid) const {
    double out = 0.0;
    int st = 0;
    pars_invalid = false;  /* never set by the reference either (:911-916) */
    getLogLikeBatch(p.data(), 1, p.getNParams(), &out, &st);
    if (st && st != POFD_E_BADLIKE)
      pofd_amd_detail::throw_status(ctx.get(), st, "calcLike", "getLogLike");
    return out;
  }

  /* many walkers per launch: theta rows [knots..., sigma_mult, ...] */
  void getLogLikeBatch(const float* theta, unsigned int nwalk, unsigned int stride,
                       double* loglike, int* status) const {
    const double half_log_2pi = 0.918938517570495605469;
    if (sets.empty()) {
      for (unsigned w = 0; w < nwalk; ++w) loglike[w] = NAN;
      return;
    }
    if (stride < 2) throw affineExcept("calcLike", "getLogLike", "Not enough elements in params");
    const unsigned nk = model.getNKnots();
    if (nk == 0) throw affineExcept("calcLike", "getLogLike", "Model knot positions not loaded");
    if (nk > stride - 1) throw affineExcept("calcLike", "getLogLike", "Not enough elements in paramSet");
    if (!prepared) throw affineExcept("calcLikeSingle", "getNegLogLike", "R range not set");
    int rc = pofd1d_loglike_batch(ctx.get(), theta, nwalk, stride, loglike,
                                  reinterpret_cast<int32_t*>(status));
    if (rc) pofd_amd_detail::throw_status(ctx.get(), rc, "calcLike", "getLogLike");
    for (unsigned w = 0; w < nwalk; ++w) {
      if (status[w] != 0) continue;
      const float* t = theta + static_cast<size_t>(w) * stride;
      const double sigmult = t[nk];
      if (has_sigma_prior) {
        const double val = (sigmult - 1.0) / sigma_prior_width;
        loglike[w] -= half_log_2pi + log(sigma_prior_width) + 0.5 * val * val;
      }
      if (regularization_alpha > 0.0) {
        numberCountsKnotsSpline m = model;
        m.setParams(paramSet(stride, t));
        loglike[w] += m.differenceRegularize(regularization_alpha);
      }
    }
  }
};

#endif
