/*
 * pofd_affine_amd.hpp -- the reference's C++ likelihood API over the C ABI.
 *
 * Header-only layer giving the classes the reference's MPI slaves and CLIs
 * call -- affineExcept, paramSet, beam, fitsData, numberCountsKnotsSpline,
 * PD, PDFactory and calcLike, and for two bands doublebeam, fitsDataDouble,
 * numberCountsDoubleLogNormal, PDDouble, PDFactoryDouble and calcLikeDouble
 * -- with the reference's names, signatures and error behaviour, backed by
 * libpofd_affine_amd.so on a gfx950 GPU.
 * Reference interfaces: include/affineExcept.h, include/paramSet.h:17-71,
 * include/beam.h:28-175, include/fitsData.h:16-96,
 * include/numberCountsKnotsSpline.h:21-70, include/PD.h:24-96,
 * include/PDFactory.h:26-153, include/calcLike.h:193-287,
 * include/doublebeam.h:78-150, include/fitsDataDouble.h:44-100,
 * include/numberCountsDoubleLogNormal.h:133-245, include/PDDouble.h:30-100,
 * include/PDFactoryDouble.h:120-160, include/calcLikeDouble.h:239-360.
 *
 * Differences a caller can observe:
 *  - calcLike::getLogLikeBatch evaluates many walkers in one launch;
 *  - the bonus quantities behind the CFIRB and Poisson priors and
 *    fillBonusParams (mean flux and flux^2 per area) are integrated to
 *    double precision on the device; the reference's GSL QAG-41 stops at
 *    relative 1e-5;
 *  - PDFactoryDouble keeps a pointer to the doublebeam given to initPD,
 *    which must outlive it;
 *  - PDFactory / PDFactoryDouble::initPD accept the abstract numberCounts /
 *    numberCountsDouble, but the device path evaluates only the models this
 *    header defines (numberCountsKnotsSpline, numberCountsDoubleLogNormal);
 *    another derived model is rejected with an affineExcept;
 *  - PD::getLogLike / PDDouble::getLogLike run on the GPU of the calling
 *    thread's default device, through a context kept per thread.
 */
#ifndef POFD_AFFINE_AMD_HPP
#define POFD_AFFINE_AMD_HPP

#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
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

/* HDF5 output (writeToHDF5 / writeToHDF5Handle) where the HDF5 header is on
   the include path, unless POFD_AFFINE_NO_HDF5 is defined */
#if !defined(POFD_AFFINE_NO_HDF5) && defined(__has_include)
#if __has_include(<hdf5.h>)
#define POFD_AFFINE_HAVE_HDF5 1
#include "pofd_affine_amd_hdf5.hpp"
#endif
#endif

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
  /* the same beam data: copies of one beam share it, and a histogram made
     since this copy was taken changes the summary it holds */
  bool sameAs(const beam& o) const {
    return h == o.h && std::memcmp(info_, o.info_, sizeof info_) == 0;
  }
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

/* include/numberCounts.h:21-70: the abstract one-band counts model (the part
   of its interface the device path needs) */
class numberCounts {
 public:
  virtual ~numberCounts() {}
  virtual void setParams(const paramSet& params) = 0;
  virtual bool isValid() const noexcept = 0;
  virtual double getMinFlux() const noexcept = 0;
  virtual double getMaxFlux() const noexcept = 0;
  virtual dblpair getRRange(const beam&) const = 0;
  /* numberCounts::differenceRegularize: no penalty unless a model has one */
  virtual double differenceRegularize(double) const { return 0.0; }
};

/* include/numberCountsKnotsSpline.h: knot positions + float knot values */
class numberCountsKnotsSpline : public numberCounts {
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
  void setParams(const paramSet& F) override {
    if (knots.size() > F.getNParams())
      throw affineExcept("numberCountsKnotsSpline", "setKnots",
                         "Number of knot values different than expected");
    values.assign(F.data(), F.data() + knots.size());
    loaded = true;
  }
  /* numberCountsKnots::isValid, src/numberCountsKnots.cc:273-284 */
  bool isValid() const noexcept override {
    if (knots.empty() || !loaded) return false;
    for (float v : values)
      if (std::isnan(v)) return false;
    return true;
  }
  unsigned int getNKnots() const noexcept { return static_cast<unsigned>(knots.size()); }
  double getKnotPos(unsigned int i) const noexcept { return knots[i]; }
  double getMinFlux() const noexcept override { return knots.empty() ? NAN : knots.front(); }
  double getMaxFlux() const noexcept override { return knots.empty() ? NAN : knots.back(); }
  const std::vector<double>& positions() const { return knots; }
  const std::vector<float>& knotValues() const { return values; }
  /* numberCountsKnots::getRRangeInternal, src/numberCountsKnots.cc:345-365 */
  dblpair getRRange(const beam& bm) const override {
    if (!isValid()) throw affineExcept("numberCountsKnots", "getRRange", "Model is not valid");
    if (!bm.hasData()) throw affineExcept("numberCountsKnots", "getRRange", "Beam is empty");
    const double maxknot = knots.back();
    if (bm.hasPos())
      return dblpair(bm.hasNeg() ? -maxknot * bm.getMinMaxNeg().second : 0.0,
                     maxknot * bm.getMinMaxPos().second);
    return dblpair(-maxknot * bm.getMinMaxNeg().second, 0.0);
  }
  /* numberCountsKnotsSpline::differenceRegularize, src/numberCountsKnotsSpline.cc:540-559 */
  double differenceRegularize(double alpha) const override {
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
  /* PD::getLogLike, include/PD.h:86, src/PD.cc:538-677 (on the device) */
  double getLogLike(const fitsData& data) const;
#ifdef POFD_AFFINE_HAVE_HDF5
  /* PD::writeToHDF5Handle / writeToHDF5, src/PD.cc:489-526: IsLog, dFlux
     and MinFlux attributes, Flux and PD datasets */
  void writeToHDF5Handle(hid_t objid) const {
    pofd_hdf5::need(objid >= 0 && H5Iget_ref(objid) >= 0, "PD", "writeToHDF5Handle",
                    "Output handle not open");
    pofd_hdf5::writeAttBool(objid, "IsLog", logflat);
    pofd_hdf5::writeAttDouble(objid, "dFlux", dflux);
    pofd_hdf5::writeAttDouble(objid, "MinFlux", minflux);
    std::vector<double> flux(getDim());
    for (unsigned i = 0; i < getDim(); ++i) flux[i] = static_cast<double>(i) * dflux + minflux;
    pofd_hdf5::writeDataDoubles(objid, "Flux", flux);
    pofd_hdf5::writeDataDoubles(objid, "PD", store);
  }
  void writeToHDF5(const std::string& outputfile) const {
    pofd_hdf5::withNewFile(outputfile, "PD", "writeToHDF5",
                           [&](hid_t f) { writeToHDF5Handle(f); });
  }
#endif
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
/* the calling thread's context for stand-alone device work (contexts are
   bound to one host thread) */
inline pofd_ctx* thread_ctx(const char* cls, const char* meth) {
  thread_local CtxPtr c;
  if (!c) c = make_ctx(default_device(), cls, meth);
  return c.get();
}
}  // namespace pofd_amd_detail

inline double PD::getLogLike(const fitsData& data) const {
  if (pd_ == nullptr || getDim() == 0)
    throw affineExcept("PD", "getLogLike", "pd not filled before likelihood calc");
  if (data.getN() == 0) throw affineExcept("PD", "getLogLike", "No data present");
  pofd_ctx* c = pofd_amd_detail::thread_ctx("PD", "getLogLike");
  double out = 0.0;
  int rc;
  if (data.isBinned())
    rc = pofd_pd_loglike(c, pd_, getDim(), minflux, dflux, logflat ? 1 : 0, nullptr, 0,
                         reinterpret_cast<const uint32_t*>(data.getBinnedData()), data.getNBins(),
                         data.getBinCent0(), data.getBinDelta(), &out);
  else
    rc = pofd_pd_loglike(c, pd_, getDim(), minflux, dflux, logflat ? 1 : 0, data.getData(),
                         data.getN(), nullptr, 0, 0.0, 0.0, &out);
  if (rc) pofd_amd_detail::throw_status(c, rc, "PD", "getLogLike");
  return out;
}

/* include/PDFactory.h: initPD + getPD for one model, one noise level.
   The device context is built on the first getPD and kept while the knot
   positions, R range, beam and sigma stay those it was built for: repeated
   getPD calls (new knot values through initPD, the same sigma) only run the
   batch kernels. */
class PDFactory {
 private:
  unsigned int ninterp;
  unsigned int n = 0;
  double minflux = 0.0, maxflux = 0.0;
  numberCountsKnotsSpline model;
  beam bm;
  bool initialized = false;
  int device = pofd_amd_detail::default_device();
  pofd_amd_detail::CtxPtr ctx;
  bool stale = true;  // the context predates a set-up change
  double built_sigma = NAN;
  int bid = 0, fid = 0;

  void build(double sigma) {
    using namespace pofd_amd_detail;
    ctx.reset();
    CtxPtr c = make_ctx(device, "PDFactory", "getPD");
    const std::vector<double>& k = model.positions();
    int rc = pofd1d_configure(c.get(), n, ninterp, k.data(), static_cast<uint32_t>(k.size()), 0.0);
    if (!rc) rc = pofd1d_add_beamset(c.get(), bm.handle(), &bid);
    const uint32_t one = 1;
    if (!rc) rc = pofd1d_add_field(c.get(), bid, nullptr, 1, &one, 1, 0.0, 1.0, sigma, 1.0, &fid);
    if (!rc) rc = pofd1d_prepare(c.get());
    if (!rc) rc = pofd1d_set_rrange(c.get(), bid, minflux, maxflux);
    if (rc) throw_status(c.get(), rc, "PDFactory", "getPD");
    ctx = std::move(c);
    built_sigma = sigma;
    stale = false;
  }

 public:
  explicit PDFactory(unsigned int NINTERP = 2048) : ninterp(NINTERP) {}
  unsigned int getNInterp() const { return ninterp; }
  void setNInterp(unsigned int v) {
    if (v != ninterp) stale = true;
    ninterp = v;
  }
  void setDevice(int d) {
    if (d != device) stale = true;
    device = d;
  }
  /* PDFactory::initPD, src/PDFactory.cc:1031-1141 (checks as written) */
  bool initPD(unsigned int N, double MINFLUX, double MAXFLUX, const numberCounts& m,
              const beam& b) {
    if (N == 0) throw affineExcept("PDFactory", "initPD", "Invalid (zero) n");
    if (N == 1) throw affineExcept("PDFactory", "initPD", "Invalid (==0) n");
    if (MAXFLUX <= 0.0) throw affineExcept("PDFactory", "initPD", "Invalid (non-positive) maxflux");
    if (MINFLUX == MAXFLUX) throw affineExcept("PDFactory", "initPD", "Invalid (0) range between min/max flux");
    if (!m.isValid()) throw affineExcept("PDFactory", "initPD", "Invalid model");
    if (!b.hasData()) throw affineExcept("PDFactory", "initPD", "Empty beam");
    const numberCountsKnotsSpline* ks = dynamic_cast<const numberCountsKnotsSpline*>(&m);
    if (!ks)
      throw affineExcept("PDFactory", "initPD",
                         "the device path evaluates numberCountsKnotsSpline models only");
    // the knot values travel with every batch; anything else rebuilds
    stale = stale || N != n || MINFLUX != minflux || MAXFLUX != maxflux ||
            ks->positions() != model.positions() || !bm.sameAs(b);
    n = N;
    minflux = MINFLUX;
    maxflux = MAXFLUX;
    model = *ks;
    bm = b;
    initialized = true;
    return true;
  }
  /* PDFactory::getPD, src/PDFactory.cc:1154-1243 */
  void getPD(double sigma, PD& pd, bool setLog = true) {
    if (!initialized) throw affineExcept("PDFactory", "getPD", "Must call initPD first");
    using namespace pofd_amd_detail;
    if (stale || !ctx || !(sigma == built_sigma)) build(sigma);
    std::vector<float> th(model.knotValues());
    th.push_back(1.0f);
    double ll = 0.0;
    int32_t st = 0;
    int rc = pofd1d_loglike_batch(ctx.get(), th.data(), 1, static_cast<uint32_t>(th.size()), &ll, &st);
    if (rc) throw_status(ctx.get(), rc, "PDFactory", "getPD");
    if (st) throw_status(ctx.get(), st, "PDFactory", "getPD");
    pd.resize(n);
    double mf = 0.0, df = 0.0;
    rc = pofd1d_get_pd(ctx.get(), 0, bid, fid, setLog ? 1 : 0, pd.pd_, &mf, &df);
    if (rc) throw_status(ctx.get(), rc, "PDFactory", "getPD");
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
  std::vector<std::string> set_beamfiles;                   // per beam set
  std::vector<double> sigmas, like_norms;
  double exp_conf = 0.0;
  bool has_sigma_prior = false;
  double sigma_prior_width = 0.0;
  bool has_cfirb_prior = false, has_poisson_prior = false;
  double cfirb_prior_mean = 0.0, cfirb_prior_sigma = 0.0;
  double poisson_prior_mean = 0.0, poisson_prior_sigma = 0.0;
  double regularization_alpha = 0.0;
  int device = pofd_amd_detail::default_device();
  mutable pofd_amd_detail::CtxPtr ctx;
  mutable bool prepared = false;
  mutable paramSet meanParams;  // parameters of the stored bonus values
  mutable double mean_flux_per_area = NAN, mean_fluxsq_per_area = NAN;

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
    set_beamfiles.clear();
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
      set_beamfiles.push_back(g.first);
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
  /* calcLike::setCFIRBPrior / setPoissonPrior, src/calcLike.cc:839-867 */
  void setCFIRBPrior(double mn, double sg) {
    has_cfirb_prior = true;
    if (mn < 0.0) throw affineExcept("calcLike", "setCFIRBPrior", "Invalid (negative) CFIRB mean");
    if (sg < 0.0) throw affineExcept("calcLike", "setCFIRBPrior", "Invalid (negative) CFIRB sigma");
    cfirb_prior_mean = mn;
    cfirb_prior_sigma = sg;
  }
  void unsetCFIRBPrior() noexcept { has_cfirb_prior = false; }
  void setPoissonPrior(double mn, double sg) {
    has_poisson_prior = true;
    if (mn < 0.0)
      throw affineExcept("calcLike", "setPoissonPrior", "Invalid (negative) Poisson prior mean");
    if (sg < 0.0)
      throw affineExcept("calcLike", "setPoissonPrior", "Invalid (negative) Poisson prior sigma");
    poisson_prior_mean = mn;
    poisson_prior_sigma = sg;
  }
  void unsetPoissonPrior() noexcept { has_poisson_prior = false; }
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
    double bonus[3];
    getLogLikeBatch(p.data(), 1, p.getNParams(), &out, &st, bonus);
    if (st && st != POFD_E_BADLIKE)
      pofd_amd_detail::throw_status(ctx.get(), st, "calcLike", "getLogLike");
    meanParams = p;
    mean_flux_per_area = bonus[1];
    mean_fluxsq_per_area = bonus[2];
    return out;
  }

  /* calcLike::fillBonusParams, src/calcLike.cc:958-980: mean flux and
     flux^2 per area into par[nknots+1], par[nknots+2]; recomputed (one
     device evaluation) when par is not the last evaluated set */
  void fillBonusParams(paramSet& par) const {
    const unsigned nk = model.getNKnots();
    if (par.getNParams() < nk + 3)
      throw affineExcept("calcLike", "fillBonusParams", "Not enough room to fill");
    float lensum = 0.0f, diffsum = 0.0f;
    const bool have = meanParams.getNParams() >= nk;
    for (unsigned i = 0; have && i < nk; ++i) {
      const float v = meanParams[i], d = v - par[i];
      lensum += v * v;
      diffsum += d * d;
    }
    const float dist = !have ? 1.0f : (lensum == 0.0f ? std::sqrt(diffsum) : std::sqrt(diffsum / lensum));
    if (dist > 1e-4f) {
      double ll = 0.0, bonus[3];
      int st = 0;
      getLogLikeBatch(par.data(), 1, par.getNParams(), &ll, &st, bonus);
      meanParams = par;
      mean_flux_per_area = bonus[1];
      mean_fluxsq_per_area = bonus[2];
    }
    par[nk + 1] = static_cast<float>(mean_flux_per_area);
    par[nk + 2] = static_cast<float>(mean_fluxsq_per_area);
  }

  /* many walkers per launch: theta rows [knots..., sigma_mult, ...];
     bonus (optional, 3 per walker): getNS, getFluxPerArea, getFluxSqPerArea */
  void getLogLikeBatch(const float* theta, unsigned int nwalk, unsigned int stride,
                       double* loglike, int* status, double* bonus = nullptr) const {
    if (sets.empty()) {
      for (unsigned w = 0; w < nwalk; ++w) loglike[w] = NAN;
      return;
    }
    if (stride < 2) throw affineExcept("calcLike", "getLogLike", "Not enough elements in params");
    const unsigned nk = model.getNKnots();
    if (nk == 0) throw affineExcept("calcLike", "getLogLike", "Model knot positions not loaded");
    if (nk > stride - 1) throw affineExcept("calcLike", "getLogLike", "Not enough elements in paramSet");
    if (!prepared) throw affineExcept("calcLikeSingle", "getNegLogLike", "R range not set");
    // the priors (src/calcLike.cc:923-952) are added on the device after
    // the dataset sum, for walkers whose datasets did not fail
    int rc = pushPriors();
    if (!rc)
      rc = pofd1d_loglike_batch(ctx.get(), theta, nwalk, stride, loglike,
                                reinterpret_cast<int32_t*>(status));
    if (rc) pofd_amd_detail::throw_status(ctx.get(), rc, "calcLike", "getLogLike");
    if (bonus) {
      rc = pofd1d_get_bonus(ctx.get(), nwalk, bonus);
      if (rc) pofd_amd_detail::throw_status(ctx.get(), rc, "calcLike", "getLogLike");
    }
  }

  /* the prior settings onto the context (pofd1d_set_priors) */
  int pushPriors() const {
    if (!ctx) return POFD_E_ARG;
    int flags = 0;
    if (has_sigma_prior) flags |= POFD_PRIOR_SIGMA;
    if (has_cfirb_prior) flags |= POFD_PRIOR_CFIRB;
    if (has_poisson_prior) flags |= POFD_PRIOR_POISSON;
    return pofd1d_set_priors(ctx.get(), flags, sigma_prior_width, cfirb_prior_mean,
                             cfirb_prior_sigma, poisson_prior_mean, poisson_prior_sigma,
                             regularization_alpha);
  }

#ifdef POFD_AFFINE_HAVE_HDF5
  /* calcLike::writeToHDF5Handle, src/calcLike.cc:986-1096: the transform
     and prior settings as attributes, the knot positions (and the last
     parameters' log10 knot values) in a Model group, and per beam set a
     BeamSets/BeamSet<i> group (calcLikeSingle::writeToHDF5Handle,
     :449-509) with NFiles, Filenames, BeamFilename and the data and R flux
     ranges (NaN before setRRanges).  The data and mask FITS extensions the
     reference also records are not kept here. */
  void writeToHDF5Handle(hid_t objid) const {
    using namespace pofd_hdf5;
    need(objid >= 0 && H5Iget_ref(objid) >= 0, "calcLike", "writeToHDF5",
         "Input handle is not valid");
    writeAttUnsignedInt(objid, "FFTSize", fftsize);
    writeAttUnsignedInt(objid, "NRInterp", ninterp);
    writeAttUnsignedInt(objid, "NBeamSets", getNBeamSets());
    writeAttBool(objid, "WasDataBinned", bin_data);
    if (bin_data) writeAttUnsignedInt(objid, "NDataBins", nbins);
    writeAttBool(objid, "HasCfirbPrior", has_cfirb_prior);
    if (has_cfirb_prior) {
      writeAttDouble(objid, "CfirbPriorMean", cfirb_prior_mean);
      writeAttDouble(objid, "CfirbPriorSigma", cfirb_prior_sigma);
    }
    writeAttBool(objid, "HasPoissonPrior", has_poisson_prior);
    if (has_poisson_prior) {
      writeAttDouble(objid, "PoissonPriorMean", poisson_prior_mean);
      writeAttDouble(objid, "PoissonPriorSigma", poisson_prior_sigma);
    }
    writeAttBool(objid, "HasSigmaPrior", has_sigma_prior);
    if (has_sigma_prior) writeAttDouble(objid, "SigmaPriorWidth", sigma_prior_width);
    writeAttBool(objid, "DidRegularize", regularization_alpha > 0.0);
    if (regularization_alpha > 0.0)
      writeAttDouble(objid, "RegularizationAlpha", regularization_alpha);
    {
      Id g(createGroup(objid, "Model", "calcLike", "writeToHDF5Handle"), H5Gclose);
      writeAttUnsignedInt(g.get(), "NKnots", model.getNKnots());
      if (model.getNKnots() > 0) {
        writeDataDoubles(g.get(), "KnotPositions", model.positions());
        const std::vector<float>& kv = model.knotValues();
        if (kv.size() >= model.getNKnots())
          writeDataDoubles(g.get(), "Log10KnotValues",
                           std::vector<double>(kv.begin(), kv.begin() + model.getNKnots()));
      }
    }
    Id bsg(createGroup(objid, "BeamSets", "calcLike", "writeToHDF5Handle"), H5Gclose);
    for (size_t s = 0; s < sets.size(); ++s) {
      Id g(createGroup(bsg.get(), "BeamSet" + std::to_string(s), "calcLikeSingle",
                       "writeToNewHDF5Group"),
           H5Gclose);
      const std::vector<size_t>& idx = sets[s].second;
      writeAttUnsignedInt(g.get(), "NFiles", static_cast<unsigned>(idx.size()));
      std::vector<std::string> names;
      for (size_t i : idx) names.push_back(datasets[i].getFile());
      if (!names.empty() && !names[0].empty()) writeDataStrings(g.get(), "Filenames", names);
      if (s < set_beamfiles.size() && !set_beamfiles[s].empty())
        writeDataStrings(g.get(), "BeamFilename", {set_beamfiles[s]});
      double info[8];
      for (double& v : info) v = NAN;
      if (prepared) (void)pofd1d_beamset_info(ctx.get(), static_cast<int>(s), info);
      writeDataDouble(g.get(), "MinDataFlux", info[3]);
      writeDataDouble(g.get(), "MaxDataFlux", info[4]);
      writeDataDouble(g.get(), "MinRFlux", info[0]);
      writeDataDouble(g.get(), "MaxRFlux", info[1]);
    }
  }
#endif

  /* the device context (after setRRanges), for the sampler: its batches
     include the priors set here */
  pofd_ctx* context() const {
    if (!prepared) throw affineExcept("calcLikeSingle", "getNegLogLike", "R range not set");
    const int rc = pushPriors();
    if (rc) pofd_amd_detail::throw_status(ctx.get(), rc, "calcLike", "context");
    return ctx.get();
  }
};

/* ======================================================================
   Two-band classes: doublebeam, fitsDataDouble, numberCountsDoubleLogNormal,
   PDDouble, PDFactoryDouble, calcLikeDouble (include/doublebeam.h,
   include/fitsDataDouble.h, include/numberCountsDoubleLogNormal.h,
   include/PDDouble.h, include/PDFactoryDouble.h, include/calcLikeDouble.h)
   ====================================================================== */

/* include/doublebeam.h: two beams split into the pp, pn, np, nn sign
   components, optionally histogrammed (host, bit-exact) */
class doublebeam {
 private:
  std::shared_ptr<pofd_dbeam> h;
  double info_[64];
  void refresh() {
    if (h) pofd_dbeam_info(h.get(), info_);
  }
  static void read_one(const std::string& f, std::vector<double>& pix, double& ps) {
    uint64_t n = 0;
    if (pofd_fits_read_beam(f.c_str(), nullptr, 0, &n, &ps))
      throw affineExcept("doublebeam", "readFiles", pofd_fits_last_error());
    pix.assign(n, 0.0);
    if (pofd_fits_read_beam(f.c_str(), pix.data(), n, &n, &ps))
      throw affineExcept("doublebeam", "readFiles", pofd_fits_last_error());
  }

 public:
  doublebeam() { std::memset(info_, 0, sizeof(info_)); }
  doublebeam(const std::string& f1, const std::string& f2, bool histogram = false,
             unsigned int NBINS = 150, double MINVAL = 1e-6) {
    std::memset(info_, 0, sizeof(info_));
    readFiles(f1, f2, MINVAL);
    if (histogram) makeHistogram(NBINS);
  }
  /* doublebeam::readFiles, src/doublebeam.cc:126-330 */
  void readFiles(const std::string& f1, const std::string& f2, double MINVAL = 1e-6) {
    std::vector<double> p1, p2;
    double ps1 = 0.0, ps2 = 0.0;
    read_one(f1, p1, ps1);
    read_one(f2, p2, ps2);
    if (p1.size() != p2.size())
      throw affineExcept("doublebeam", "readFiles", "Beam sizes don't match");
    if (ps1 != ps2)
      throw affineExcept("doublebeam", "readFiles", "Beam pixel sizes don't match");
    setBeams(static_cast<unsigned>(p1.size()), p1.data(), p2.data(), ps1, MINVAL);
  }
  /* doublebeam::setBeams, src/doublebeam.cc:331-466 */
  void setBeams(unsigned int n, const double* const b1, const double* const b2,
                double PIXSIZE, double MINVAL = 1e-6) {
    pofd_dbeam* b = nullptr;
    if (pofd_dbeam_create(&b, b1, b2, n, PIXSIZE, MINVAL))
      throw affineExcept("doublebeam", "setBeams", pofd_dbeam_last_error());
    h.reset(b, pofd_dbeam_destroy);
    refresh();
  }
  /* doublebeam::makeHistogram, src/doublebeam.cc:467-584 */
  void makeHistogram(unsigned int NBINS = 150) {
    if (!h) throw affineExcept("doublebeam", "makeHistogram", "No beam read");
    if (pofd_dbeam_make_histogram(h.get(), NBINS))
      throw affineExcept("doublebeam", "makeHistogram", pofd_dbeam_last_error());
    refresh();
  }
  bool hasData() const {
    for (int s = 0; s < 4; ++s)
      if (hasSign(s)) return true;
    return false;
  }
  bool hasSign(unsigned int i) const { return info_[12 * i] != 0.0; }
  unsigned int getNPix(unsigned int i) const { return static_cast<unsigned>(info_[12 * i + 1]); }
  unsigned int getMaxNPix() const {
    unsigned m = 0;
    for (int s = 0; s < 4; ++s) m = std::max(m, getNPix(s));
    return m;
  }
  unsigned int getTotalNPix() const {
    unsigned t = 0;
    for (int s = 0; s < 4; ++s) t += getNPix(s);
    return t;
  }
  bool isHistogrammed(unsigned int i) const { return info_[12 * i + 2] != 0.0; }
  unsigned int getNHist(unsigned int i) const { return static_cast<unsigned>(info_[12 * i + 3]); }
  dblpair getMinMax1(unsigned int i) const { return dblpair(info_[12 * i + 8], info_[12 * i + 9]); }
  dblpair getMinMax2(unsigned int i) const { return dblpair(info_[12 * i + 10], info_[12 * i + 11]); }
  double getPixSize() const { return info_[48]; }
  double getMinval() const { return info_[49]; }
  unsigned int getNBins() const { return static_cast<unsigned>(info_[50]); }
  /* per-sign effective areas in pixels (getEffectiveAreaSign1/2) */
  double getEffectiveAreaSign1(unsigned int i) const { return info_[12 * i + 4]; }
  double getEffectiveAreaSign2(unsigned int i) const { return info_[12 * i + 5]; }
  double getEffectiveAreaSqSign1(unsigned int i) const { return info_[12 * i + 6]; }
  double getEffectiveAreaSqSign2(unsigned int i) const { return info_[12 * i + 7]; }
  /* effective areas in square degrees */
  double getEffectiveArea1() const {
    double a = 0.0;
    for (int s = 0; s < 4; ++s) a += getEffectiveAreaSign1(s);
    const double ps = getPixSize() / 3600.0;
    return a * ps * ps;
  }
  double getEffectiveArea2() const {
    double a = 0.0;
    for (int s = 0; s < 4; ++s) a += getEffectiveAreaSign2(s);
    const double ps = getPixSize() / 3600.0;
    return a * ps * ps;
  }
  const pofd_dbeam* handle() const { return h.get(); }
};

/* include/fitsDataDouble.h: pixel pairs of two maps, optionally binned
   (host, bit-exact) */
class fitsDataDouble {
 private:
  std::vector<double> d1, d2;
  bool is_binned = false;
  unsigned int nb1 = 0, nb2 = 0;
  double bininfo[4] = {0.0, 0.0, 0.0, 0.0};
  std::vector<uint32_t> binval;
  std::string file1, file2;

 public:
  fitsDataDouble() {}
  fitsDataDouble(const std::string& f1, const std::string& f2, bool ignore_mask = false,
                 bool meansub = false) {
    readData(f1, f2, ignore_mask, meansub);
  }
  /* fitsDataDouble::readData, src/fitsDataDouble.cc:305-406 */
  void readData(const std::string& f1, const std::string& f2, bool ignore_mask = false,
                bool meansub = false) {
    uint64_t n = 0;
    const int ig = ignore_mask ? 1 : 0;
    if (pofd_fits_read_map_pair(f1.c_str(), f2.c_str(), ig, nullptr, nullptr, 0, &n))
      throw affineExcept("fitsDataDouble", "readData", pofd_fits_last_error());
    d1.assign(n, 0.0);
    d2.assign(n, 0.0);
    if (pofd_fits_read_map_pair(f1.c_str(), f2.c_str(), ig, d1.data(), d2.data(), n, &n))
      throw affineExcept("fitsDataDouble", "readData", pofd_fits_last_error());
    file1 = f1;
    file2 = f2;
    removeBinning();
    if (meansub) meanSubtract();
  }
  void setData(uint64_t n, const double* a, const double* b) {
    d1.assign(a, a + n);
    d2.assign(b, b + n);
    removeBinning();
  }
  bool isBinned() const { return is_binned; }
  /* fitsDataDouble::applyBinning, src/fitsDataDouble.cc:517-566 */
  void applyBinning(unsigned int NB1, unsigned int NB2) {
    if (d1.empty()) return;
    if (NB1 == 0 || NB2 == 0)
      throw affineExcept("fitsDataDouble", "applyBinning", "Trying to bin with no bins");
    if (is_binned && NB1 == nb1 && NB2 == nb2) return;
    binval.assign(static_cast<size_t>(NB1) * NB2, 0u);
    if (pofd_data2_bin(d1.data(), d2.data(), d1.size(), NB1, NB2, binval.data(), bininfo))
      throw affineExcept("fitsDataDouble", "applyBinning", "binning failed");
    nb1 = NB1;
    nb2 = NB2;
    is_binned = true;
  }
  void removeBinning() {
    is_binned = false;
    nb1 = nb2 = 0;
    binval.clear();
  }
  std::pair<unsigned int, unsigned int> getNBins() const { return std::make_pair(nb1, nb2); }
  dblpair getBinCent0() const { return dblpair(bininfo[0], bininfo[2]); }
  dblpair getBinDelta() const { return dblpair(bininfo[1], bininfo[3]); }
  const double* binInfo() const { return bininfo; }
  const unsigned int* getBinnedData() const {
    return reinterpret_cast<const unsigned int*>(binval.data());
  }
  bool hasData() const { return !d1.empty(); }
  unsigned int getN() const { return static_cast<unsigned>(d1.size()); }
  const double* getData1() const { return d1.data(); }
  const double* getData2() const { return d2.data(); }
  /* fitsDataDouble::meanSubtract, src/fitsDataDouble.cc:567-608 */
  dblpair meanSubtract() {
    double m[2] = {0.0, 0.0};
    if (!d1.empty()) pofd_data2_mean_subtract(d1.data(), d2.data(), d1.size(), m);
    if (is_binned) {
      const unsigned a = nb1, b = nb2;
      removeBinning();
      applyBinning(a, b);
    }
    return dblpair(m[0], m[1]);
  }
  std::pair<dblpair, dblpair> getMinMax() const {
    if (d1.empty()) return std::make_pair(dblpair(NAN, NAN), dblpair(NAN, NAN));
    double a0 = d1[0], a1 = d1[0], b0 = d2[0], b1 = d2[0];
    for (size_t i = 1; i < d1.size(); ++i) {
      a0 = std::min(a0, d1[i]);
      a1 = std::max(a1, d1[i]);
      b0 = std::min(b0, d2[i]);
      b1 = std::max(b1, d2[i]);
    }
    return std::make_pair(dblpair(a0, b0), dblpair(a1, b1));
  }
  dblpair getMin() const { return getMinMax().first; }
  dblpair getMax() const { return getMinMax().second; }
  std::pair<std::string, std::string> getFiles() const { return std::make_pair(file1, file2); }
};

/* include/numberCountsDouble.h:22-84: the abstract two-band counts model
   (the part of its interface the device path needs) */
class numberCountsDouble {
 public:
  virtual ~numberCountsDouble() {}
  virtual bool isValid() const noexcept = 0;
  virtual unsigned int getNParams() const noexcept = 0;
  virtual void setParams(const paramSet& params) = 0;
  virtual double differenceRegularize(double) const { return 0.0; }
};

/* include/numberCountsDoubleLogNormal.h: band-1 knot spline plus the
   log-normal colour model (sigma and offset knots) */
class numberCountsDoubleLogNormal : public numberCountsDouble {
 private:
  std::vector<double> knots, sigmas, offsets;
  std::vector<float> values;  // knot values, sigma values, offset values
  double min_sigma = 0.0;
  bool loaded = false;
  static void check_pos(const std::vector<double>& v, const char* meth) {
    for (size_t i = 0; i < v.size(); ++i) {
      if (v[i] <= 0.0)
        throw affineExcept("numberCountsDoubleLogNormal", meth, "Negative knot positions not allowed");
      if (i > 0 && v[i] <= v[i - 1])
        throw affineExcept("numberCountsDoubleLogNormal", meth, "Knot positions not monotonically increasing");
    }
  }

 public:
  numberCountsDoubleLogNormal() {}
  numberCountsDoubleLogNormal(const std::vector<double>& K, const std::vector<double>& S,
                              const std::vector<double>& O) {
    setPositions(K, S, O);
  }
  /* setKnotPositions / setSigmaPositions / setOffsetPositions,
     src/numberCountsDoubleLogNormal.cc:632-920 */
  void setKnotPositions(const std::vector<double>& K) {
    check_pos(K, "setKnotPositions");
    knots = K;
    loaded = false;
  }
  void setSigmaPositions(const std::vector<double>& S) {
    check_pos(S, "setSigmaPositions");
    sigmas = S;
    loaded = false;
  }
  void setOffsetPositions(const std::vector<double>& O) {
    check_pos(O, "setOffsetPositions");
    offsets = O;
    loaded = false;
  }
  void setPositions(const std::vector<double>& K, const std::vector<double>& S,
                    const std::vector<double>& O) {
    setKnotPositions(K);
    setSigmaPositions(S);
    setOffsetPositions(O);
  }
  /* numberCountsDoubleLogNormal::setMinSigma, :792 */
  void setMinSigma(double v) {
    if (v < 0.0) throw affineExcept("numberCountsDoubleLogNormal", "setMinSigma", "Invalid (negative) min sigma");
    min_sigma = v;
  }
  double getMinSigma() const noexcept { return min_sigma; }
  unsigned int getNKnots() const noexcept { return static_cast<unsigned>(knots.size()); }
  unsigned int getNSigmas() const noexcept { return static_cast<unsigned>(sigmas.size()); }
  unsigned int getNOffsets() const noexcept { return static_cast<unsigned>(offsets.size()); }
  unsigned int getNParams() const noexcept override { return getNKnots() + getNSigmas() + getNOffsets(); }
  double getKnotPosition(unsigned int i) const { return knots[i]; }
  double getSigmaPosition(unsigned int i) const { return sigmas[i]; }
  double getOffsetPosition(unsigned int i) const { return offsets[i]; }
  const std::vector<double>& knotPositions() const { return knots; }
  const std::vector<double>& sigmaPositions() const { return sigmas; }
  const std::vector<double>& offsetPositions() const { return offsets; }
  const std::vector<float>& paramValues() const { return values; }
  /* numberCountsDoubleLogNormal::setParams, :1010-1052 */
  void setParams(const paramSet& F) override {
    if (F.getNParams() < getNParams())
      throw affineExcept("numberCountsDoubleLogNormal", "setParams", "Not enough parameters present");
    values.assign(F.data(), F.data() + getNParams());
    loaded = true;
  }
  /* isValid: knot values not NaN, sigma values finite, offsets not NaN
     (checkKnotsValid/checkSigmasValid/checkOffsetsValid, :1053-1124) */
  bool isValid() const noexcept override {
    if (!loaded || knots.empty() || sigmas.empty() || offsets.empty()) return false;
    const size_t nk = knots.size(), ns = sigmas.size();
    for (size_t i = 0; i < values.size(); ++i) {
      const float v = values[i];
      if (std::isnan(v)) return false;
      if (i >= nk && i < nk + ns && std::isinf(v)) return false;
    }
    return true;
  }
  /* numberCountsDoubleLogNormal::differenceRegularize, :2154-2173 */
  double differenceRegularize(double alpha) const override {
    if (!isValid()) return NAN;
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

/* include/PDDouble.h: a 2D P(D), n1 x n2, band 1 the slow index */
class PDDouble {
 private:
  std::vector<double> store;

 public:
  unsigned int n1 = 0, n2 = 0;
  bool logflat = false;
  double minflux1 = 0.0, dflux1 = 1.0, minflux2 = 0.0, dflux2 = 1.0;
  double* pd_ = nullptr;
  explicit PDDouble(unsigned int N1 = 0, double MINFLUX1 = 0.0, double DFLUX1 = 1.0,
                    unsigned int N2 = 0, double MINFLUX2 = 0.0, double DFLUX2 = 1.0)
      : store(static_cast<size_t>(N1) * N2, 0.0), n1(N1), n2(N2), minflux1(MINFLUX1),
        dflux1(DFLUX1), minflux2(MINFLUX2), dflux2(DFLUX2) {
    pd_ = store.empty() ? nullptr : store.data();
  }
  void resize(unsigned int N1, unsigned int N2) {
    store.resize(static_cast<size_t>(N1) * N2);
    n1 = N1;
    n2 = N2;
    pd_ = store.empty() ? nullptr : store.data();
  }
  unsigned int getDim1() const { return n1; }
  unsigned int getDim2() const { return n2; }
  bool isLog() const { return logflat; }
  double getFluxVal1(unsigned int i) const { return minflux1 + static_cast<double>(i) * dflux1; }
  double getFluxVal2(unsigned int j) const { return minflux2 + static_cast<double>(j) * dflux2; }
  double getPDVal(unsigned int i, unsigned int j) const { return pd_[static_cast<size_t>(i) * n2 + j]; }
  /* PDDouble::getLogLike, include/PDDouble.h:95, src/PDDouble.cc:976-1222
     (on the device) */
  double getLogLike(const fitsDataDouble& data) const {
    if (pd_ == nullptr || n1 == 0 || n2 == 0)
      throw affineExcept("PDDouble", "getLogLike", "pd not filled before likelihood calc");
    if (data.getN() == 0) throw affineExcept("PDDouble", "getLogLike", "No data present");
    pofd_ctx* c = pofd_amd_detail::thread_ctx("PDDouble", "getLogLike");
    double out = 0.0;
    int rc;
    if (data.isBinned()) {
      const double* bi = data.binInfo();
      rc = pofd_pd2_loglike(c, pd_, n1, n2, minflux1, dflux1, minflux2, dflux2, logflat ? 1 : 0,
                            nullptr, nullptr, 0,
                            reinterpret_cast<const uint32_t*>(data.getBinnedData()),
                            data.getNBins().first, data.getNBins().second, bi[0], bi[1], bi[2],
                            bi[3], &out);
    } else {
      rc = pofd_pd2_loglike(c, pd_, n1, n2, minflux1, dflux1, minflux2, dflux2, logflat ? 1 : 0,
                            data.getData1(), data.getData2(), data.getN(), nullptr, 0, 0, 0.0,
                            0.0, 0.0, 0.0, &out);
    }
    if (rc) pofd_amd_detail::throw_status(c, rc, "PDDouble", "getLogLike");
    return out;
  }
#ifdef POFD_AFFINE_HAVE_HDF5
  /* PDDouble::writeToHDF5Handle / writeToHDF5, src/PDDouble.cc:912-966:
     IsLog, dFlux1, dFlux2, MinFlux1 and MinFlux2 attributes, Flux1, Flux2
     and the n1 x n2 PD datasets */
  void writeToHDF5Handle(hid_t objid) const {
    pofd_hdf5::need(objid >= 0 && H5Iget_ref(objid) >= 0, "PDDouble", "writeToHDF5Handle",
                    "Failed to open HDF5 file to write");
    pofd_hdf5::writeAttBool(objid, "IsLog", logflat);
    pofd_hdf5::writeAttDouble(objid, "dFlux1", dflux1);
    pofd_hdf5::writeAttDouble(objid, "dFlux2", dflux2);
    pofd_hdf5::writeAttDouble(objid, "MinFlux1", minflux1);
    pofd_hdf5::writeAttDouble(objid, "MinFlux2", minflux2);
    std::vector<double> f1(n1), f2(n2);
    for (unsigned i = 0; i < n1; ++i) f1[i] = static_cast<double>(i) * dflux1 + minflux1;
    for (unsigned j = 0; j < n2; ++j) f2[j] = static_cast<double>(j) * dflux2 + minflux2;
    pofd_hdf5::writeDataDoubles(objid, "Flux1", f1);
    pofd_hdf5::writeDataDoubles(objid, "Flux2", f2);
    pofd_hdf5::writeData(objid, "PD", H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, {n1, n2},
                         store.data(), "writeData2DDoubles");
  }
  void writeToHDF5(const std::string& outputfile) const {
    pofd_hdf5::withNewFile(outputfile, "PDDouble", "writeToHDF5",
                           [&](hid_t f) { writeToHDF5Handle(f); });
  }
#endif
};

namespace pofd_amd_detail {
inline void throw_status2(pofd_ctx* c, int st, const char* cls, const char* meth) {
  const std::string msg = c ? pofd_last_error(c) : "";
  switch (st) {
    case POFD_E_SPLIT:
      throw affineExcept("PDFactoryDouble", "findSplitPoint",
                         "Unable to find split point" + (msg.empty() ? std::string() : ": " + msg));
    case POFD_E_INVALID_MODEL:
      throw affineExcept("PDFactoryDouble", "initPD", "Invalid model");
    case POFD_E_MEAN:
      throw affineExcept("PDFactoryDouble", "unwrapAndNormalizePD", "Un-shift amounts not finite");
    case POFD_E_NOGPU:
      throw affineExcept(cls, meth, "no gfx950 GPU available");
    default:
      throw affineExcept(cls, meth, msg.empty() ? "device path failed" : msg);
  }
}
inline std::vector<float> theta2(const numberCountsDoubleLogNormal& m, float sm1, float sm2) {
  std::vector<float> th(m.paramValues());
  th.push_back(sm1);
  th.push_back(sm2);
  return th;
}
}  // namespace pofd_amd_detail

/* include/PDFactoryDouble.h: initPD + getPD for one model and one pair of
   noise levels.  The device context is built on the first getPD and kept
   while the model's positions, the R ranges, the beam, the edge setting
   and the two sigmas stay those it was built for. */
class PDFactoryDouble {
 private:
  unsigned int nedge;
  unsigned int n = 0;
  double range[4] = {0.0, 0.0, 0.0, 0.0};
  bool setEdge = true;
  numberCountsDoubleLogNormal model;
  const doublebeam* bm = nullptr;
  bool initialized = false;
  int device = pofd_amd_detail::default_device();
  pofd_amd_detail::CtxPtr ctx;
  bool stale = true;
  double built_sigma[2] = {NAN, NAN};
  int bid = 0, fid = 0;

  void build(double sigma1, double sigma2) {
    using namespace pofd_amd_detail;
    ctx.reset();
    CtxPtr c = make_ctx(device, "PDFactoryDouble", "getPD");
    const auto& K = model.knotPositions();
    const auto& S = model.sigmaPositions();
    const auto& O = model.offsetPositions();
    int rc = pofd2d_configure(c.get(), n, nedge, setEdge ? 1 : 0, K.data(),
                              static_cast<uint32_t>(K.size()), S.data(),
                              static_cast<uint32_t>(S.size()), O.data(),
                              static_cast<uint32_t>(O.size()), model.getMinSigma(), 0.0, 0.0);
    if (!rc) rc = pofd2d_add_beamset(c.get(), bm->handle(), &bid);
    const double z = 0.0;
    if (!rc)
      rc = pofd2d_add_field(c.get(), bid, &z, &z, 1, nullptr, 0, 0, nullptr, sigma1, sigma2, 1.0,
                            &fid);
    std::vector<float> th = theta2(model, 1.0f, 1.0f);
    if (!rc) rc = pofd2d_prepare(c.get(), th.data(), static_cast<uint32_t>(th.size()));
    if (!rc) rc = pofd2d_set_rrange(c.get(), bid, range);
    if (rc) throw_status2(c.get(), rc, "PDFactoryDouble", "getPD");
    ctx = std::move(c);
    built_sigma[0] = sigma1;
    built_sigma[1] = sigma2;
    stale = false;
  }

 public:
  explicit PDFactoryDouble(unsigned int NEDGE = 256) : nedge(NEDGE) {}
  void setNEdge(unsigned int v) {
    if (v != nedge) stale = true;
    nedge = v;
  }
  unsigned int getNEdge() const { return nedge; }
  void setDevice(int d) {
    if (d != device) stale = true;
    device = d;
  }
  /* PDFactoryDouble::initPD, src/PDFactoryDouble.cc:1143-1367 (checks as
     written); the beam must outlive the factory */
  bool initPD(unsigned int N, double minflux1, double maxflux1, double minflux2,
              double maxflux2, const numberCountsDouble& m, const doublebeam& b,
              bool setedge = true) {
    if (N == 0) throw affineExcept("PDFactoryDouble", "initPD", "Invalid (non-positive) n");
    if (maxflux1 <= 0.0 || maxflux2 <= 0.0)
      throw affineExcept("PDFactoryDouble", "initPD", "Invalid (non-positive) maxflux");
    if (!m.isValid()) throw affineExcept("PDFactoryDouble", "initPD", "Invalid model");
    if (!b.hasData()) throw affineExcept("PDFactoryDouble", "initPD", "Empty beam");
    const numberCountsDoubleLogNormal* ln = dynamic_cast<const numberCountsDoubleLogNormal*>(&m);
    if (!ln)
      throw affineExcept("PDFactoryDouble", "initPD",
                         "the device path evaluates numberCountsDoubleLogNormal models only");
    const double r[4] = {minflux1, maxflux1, minflux2, maxflux2};
    // the model's values travel with every batch (the colour model set-up
    // on the device is rebuilt from them); anything else rebuilds
    stale = stale || N != n || setedge != setEdge || &b != bm ||
            std::memcmp(r, range, sizeof r) != 0 || ln->knotPositions() != model.knotPositions() ||
            ln->sigmaPositions() != model.sigmaPositions() ||
            ln->offsetPositions() != model.offsetPositions() ||
            ln->getMinSigma() != model.getMinSigma();
    n = N;
    std::memcpy(range, r, sizeof r);
    setEdge = setedge;
    model = *ln;
    bm = &b;
    initialized = true;
    return true;
  }
  /* PDFactoryDouble::getPD, src/PDFactoryDouble.cc:1382-1499 */
  void getPD(double sigma1, double sigma2, PDDouble& pd, bool setLog = true) {
    if (!initialized) throw affineExcept("PDFactoryDouble", "getPD", "Must call initPD first");
    using namespace pofd_amd_detail;
    if (stale || !ctx || !(sigma1 == built_sigma[0]) || !(sigma2 == built_sigma[1]))
      build(sigma1, sigma2);
    std::vector<float> th = theta2(model, 1.0f, 1.0f);
    double ll = 0.0;
    int32_t st = 0;
    int rc = pofd2d_loglike_batch(ctx.get(), th.data(), 1, static_cast<uint32_t>(th.size()), &ll, &st);
    if (rc) throw_status2(ctx.get(), rc, "PDFactoryDouble", "getPD");
    if (st) throw_status2(ctx.get(), st, "PDFactoryDouble", "getPD");
    pd.resize(n, n);
    double info[8];
    rc = pofd2d_get_pd(ctx.get(), 0, bid, fid, setLog ? 1 : 0, pd.pd_, info);
    if (rc) throw_status2(ctx.get(), rc, "PDFactoryDouble", "getPD");
    pd.minflux1 = info[0];
    pd.minflux2 = info[1];
    pd.dflux1 = info[2];
    pd.dflux2 = info[3];
    pd.logflat = setLog;
  }
};

/* include/calcLikeDouble.h: data sets grouped by beam pair, priors, the
   two-band likelihood */
class calcLikeDouble {
 private:
  unsigned int fftsize, nedge;
  bool edgeInteg, bin_data;
  unsigned int nbins;
  mutable numberCountsDoubleLogNormal model;
  std::vector<fitsDataDouble> datasets;
  std::vector<std::pair<doublebeam, std::vector<size_t>>> sets;  // beam pair, dataset idx
  std::vector<double> sigmas1, sigmas2, like_norms;
  double exp_conf1 = 0.0, exp_conf2 = 0.0;
  bool has_sigma_prior1 = false, has_sigma_prior2 = false;
  double sigma_prior_width1 = 0.0, sigma_prior_width2 = 0.0;
  bool has_cfirb[2] = {false, false}, has_poisson[2] = {false, false};
  double cfirb_mean[2] = {0.0, 0.0}, cfirb_sigma[2] = {0.0, 0.0};
  double poisson_mean[2] = {0.0, 0.0}, poisson_sigma[2] = {0.0, 0.0};
  double regularization_alpha = 0.0;
  mutable paramSet meanParams;
  mutable double bonus_[5] = {NAN, NAN, NAN, NAN, NAN};
  void set_prior(bool* has, double* mn, double* sg, int band, double m, double s,
                 const char* meth, const char* what) {
    has[band] = true;
    if (m < 0.0) throw affineExcept("calcLikeDouble", meth, std::string("Invalid (negative) ") + what + " mean");
    if (s < 0.0) throw affineExcept("calcLikeDouble", meth, std::string("Invalid (negative) ") + what + " sigma");
    mn[band] = m;
    sg[band] = s;
  }
  int device = pofd_amd_detail::default_device();
  mutable pofd_amd_detail::CtxPtr ctx;
  mutable bool prepared = false;

 public:
  explicit calcLikeDouble(unsigned int FFTSIZE = 4096, unsigned int NEDGE = 256,
                          bool EDGEINTEG = true, bool BINNED = false, unsigned int NBINS = 1000)
      : fftsize(FFTSIZE), nedge(NEDGE), edgeInteg(EDGEINTEG), bin_data(BINNED), nbins(NBINS) {}
  calcLikeDouble(const calcLikeDouble&) = delete;
  calcLikeDouble& operator=(const calcLikeDouble&) = delete;

  void setDevice(int d) { device = d; prepared = false; }
  void setFFTSize(unsigned int v) noexcept { fftsize = v; prepared = false; }
  unsigned int getFFTSize() const noexcept { return fftsize; }
  void setEdgeInteg() noexcept { edgeInteg = true; prepared = false; }
  void unSetEdgeInteg() noexcept { edgeInteg = false; prepared = false; }
  bool getEdgeInteg() const noexcept { return edgeInteg; }
  void setNEdge(unsigned int v) { nedge = v; prepared = false; }
  unsigned int getNEdge() const noexcept { return nedge; }
  /* setBinData / unSetBinData / setNBins, src/calcLikeDouble.cc:1037-1092 */
  void setBinData() {
    bin_data = true;
    for (auto& d : datasets) d.applyBinning(nbins, nbins);
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
      for (auto& d : datasets) d.applyBinning(nbins, nbins);
    prepared = false;
  }

  /* calcLikeDouble::readDataFromFiles, src/calcLikeDouble.cc:931-1024:
     datasets are grouped by the pair of beam files, groups ordered by the
     key "beam1:beam2" (std::map) */
  void readDataFromFiles(const std::vector<std::string>& datafiles1,
                         const std::vector<std::string>& datafiles2,
                         const std::vector<std::string>& beamfiles1,
                         const std::vector<std::string>& beamfiles2,
                         const std::vector<double>& sigs1, const std::vector<double>& sigs2,
                         const std::vector<double>& lnorms, bool IGNOREMASK = false,
                         bool MEANSUB = false, double MINBEAMVAL = 1e-6,
                         bool HISTOGRAM = false, unsigned int NBEAMHIST = 160,
                         double EXPCONF1 = 0.0, double EXPCONF2 = 0.0) {
    const size_t ndat = datafiles1.size();
    const char* cls = "calcLikeDouble";
    const char* meth = "readDataFromFiles";
    if (ndat != datafiles2.size()) throw affineExcept(cls, meth, "Datafiles1 and datafile2 not same length");
    if (ndat == 0) throw affineExcept(cls, meth, "No datafiles");
    if (ndat != beamfiles1.size()) throw affineExcept(cls, meth, "datafiles and beamfiles1 not same length");
    if (ndat != beamfiles2.size()) throw affineExcept(cls, meth, "datafiles and beamfiles2 not same length");
    if (ndat != sigs1.size()) throw affineExcept(cls, meth, "datafiles and sigmas1 not same length");
    if (ndat != sigs2.size()) throw affineExcept(cls, meth, "datafiles and sigmas2 not same length");
    if (ndat != lnorms.size()) throw affineExcept(cls, meth, "datafiles and like_norm not same length");
    std::map<std::string, std::vector<size_t>> groups;
    for (size_t i = 0; i < ndat; ++i) groups[beamfiles1[i] + ":" + beamfiles2[i]].push_back(i);
    datasets.clear();
    sets.clear();
    sigmas1.clear();
    sigmas2.clear();
    like_norms.clear();
    for (auto& g : groups) {
      const size_t first = g.second.front();
      doublebeam bm(beamfiles1[first], beamfiles2[first], HISTOGRAM, NBEAMHIST, MINBEAMVAL);
      std::vector<size_t> idx;
      for (size_t i : g.second) {
        datasets.emplace_back();
        fitsDataDouble& d = datasets.back();
        d.readData(datafiles1[i], datafiles2[i], IGNOREMASK, MEANSUB);
        if (bin_data) d.applyBinning(nbins, nbins);
        sigmas1.push_back(sigs1[i]);
        sigmas2.push_back(sigs2[i]);
        like_norms.push_back(lnorms[i]);
        idx.push_back(datasets.size() - 1);
      }
      sets.emplace_back(std::move(bm), idx);
    }
    exp_conf1 = EXPCONF1;
    exp_conf2 = EXPCONF2;
    prepared = false;
  }

  unsigned int getNBeamSets() const noexcept { return static_cast<unsigned>(sets.size()); }
  unsigned int getNKnots() const noexcept { return model.getNKnots(); }
  unsigned int getNSigmas() const noexcept { return model.getNSigmas(); }
  unsigned int getNOffsets() const noexcept { return model.getNOffsets(); }
  void setKnotPositions(std::vector<double>& k) { model.setKnotPositions(k); prepared = false; }
  void setSigmaPositions(std::vector<double>& k) { model.setSigmaPositions(k); prepared = false; }
  void setOffsetPositions(std::vector<double>& k) { model.setOffsetPositions(k); prepared = false; }
  void setPositions(std::vector<double>& K, std::vector<double>& S, std::vector<double>& O) {
    model.setPositions(K, S, O);
    prepared = false;
  }
  void setMinSigma(double v) { model.setMinSigma(v); prepared = false; }

  /* calcLikeDouble::setRRanges, src/calcLikeDouble.cc:1093-1097: the R
     ranges depend on the model at p (the band-2 maximum flux) */
  void setRRanges(const paramSet& p) {
    using namespace pofd_amd_detail;
    model.setParams(p);
    ctx = make_ctx(device, "calcLikeDouble", "setRRanges");
    const auto& K = model.knotPositions();
    const auto& S = model.sigmaPositions();
    const auto& O = model.offsetPositions();
    int rc = pofd2d_configure(ctx.get(), fftsize, nedge, edgeInteg ? 1 : 0, K.data(),
                              static_cast<uint32_t>(K.size()), S.data(),
                              static_cast<uint32_t>(S.size()), O.data(),
                              static_cast<uint32_t>(O.size()), model.getMinSigma(), exp_conf1,
                              exp_conf2);
    for (size_t s = 0; s < sets.size() && !rc; ++s) {
      int b = 0;
      rc = pofd2d_add_beamset(ctx.get(), sets[s].first.handle(), &b);
      for (size_t i : sets[s].second) {
        if (rc) break;
        const fitsDataDouble& d = datasets[i];
        int f = 0;
        if (d.isBinned())
          rc = pofd2d_add_field(ctx.get(), b, d.getData1(), d.getData2(), d.getN(),
                                reinterpret_cast<const uint32_t*>(d.getBinnedData()),
                                d.getNBins().first, d.getNBins().second, d.binInfo(),
                                sigmas1[i], sigmas2[i], like_norms[i], &f);
        else
          rc = pofd2d_add_field(ctx.get(), b, d.getData1(), d.getData2(), d.getN(), nullptr,
                                0, 0, nullptr, sigmas1[i], sigmas2[i], like_norms[i], &f);
      }
    }
    if (!rc) rc = pofd2d_prepare(ctx.get(), p.data(), p.getNParams());
    if (rc) throw_status2(ctx.get(), rc, "calcLikeDouble", "setRRanges");
    prepared = true;
  }
  void setSigmaPrior1(double val) noexcept { has_sigma_prior1 = true; sigma_prior_width1 = val; }
  void unsetSigmaPrior1() noexcept { has_sigma_prior1 = false; }
  void setSigmaPrior2(double val) noexcept { has_sigma_prior2 = true; sigma_prior_width2 = val; }
  void unsetSigmaPrior2() noexcept { has_sigma_prior2 = false; }
  /* setCFIRBPrior1/2, setPoissonPrior1/2, src/calcLikeDouble.cc:1105-1174 */
  void setCFIRBPrior1(double m, double sg) {
    set_prior(has_cfirb, cfirb_mean, cfirb_sigma, 0, m, sg, "setCFIRBPrior1", "CFIRB");
  }
  void setCFIRBPrior2(double m, double sg) {
    set_prior(has_cfirb, cfirb_mean, cfirb_sigma, 1, m, sg, "setCFIRBPrior2", "CFIRB");
  }
  void unsetCFIRBPrior1() noexcept { has_cfirb[0] = false; }
  void unsetCFIRBPrior2() noexcept { has_cfirb[1] = false; }
  void setPoissonPrior1(double m, double sg) {
    set_prior(has_poisson, poisson_mean, poisson_sigma, 0, m, sg, "setPoissonPrior1", "Poisson prior");
  }
  void setPoissonPrior2(double m, double sg) {
    set_prior(has_poisson, poisson_mean, poisson_sigma, 1, m, sg, "setPoissonPrior2", "Poisson prior");
  }
  void unsetPoissonPrior1() noexcept { has_poisson[0] = false; }
  void unsetPoissonPrior2() noexcept { has_poisson[1] = false; }
  void setRegularizationAlpha(double alpha) {
    if (alpha < 0.0)
      throw affineExcept("calcLikeDouble", "setRegularizationAlpha", "Invalid (negative) regularization alpha");
    regularization_alpha = alpha;
  }

  /* calcLikeDouble::getLogLike, src/calcLikeDouble.cc:1195-1281 */
  double getLogLike(const paramSet& p, bool& pars_invalid) const {
    double out = 0.0;
    int st = 0;
    pars_invalid = false;  /* never set by the reference either */
    double bonus[5];
    getLogLikeBatch(p.data(), 1, p.getNParams(), &out, &st, bonus);
    if (st && st != POFD_E_BADLIKE)
      pofd_amd_detail::throw_status2(ctx.get(), st, "calcLikeDouble", "getLogLike");
    meanParams = p;
    std::copy(bonus, bonus + 5, bonus_);
    return out;
  }

  /* calcLikeDouble::fillBonusParams, src/calcLikeDouble.cc:1283-1316: mean
     flux per area of both bands, then flux^2 per area, into
     par[nmodel+2 .. nmodel+5]; recomputed (one device evaluation) when par
     is not the last evaluated set */
  void fillBonusParams(paramSet& par) const {
    const unsigned nmp = model.getNParams();
    if (par.getNParams() < nmp + 6)
      throw affineExcept("calcLikeDouble", "fillBonusParams", "Not enough room to fill");
    float lensum = 0.0f, diffsum = 0.0f;
    const bool have = meanParams.getNParams() >= nmp;
    for (unsigned i = 0; have && i < nmp; ++i) {
      const float v = meanParams[i], d = v - par[i];
      lensum += v * v;
      diffsum += d * d;
    }
    const float dist = !have ? 1.0f : (lensum == 0.0f ? std::sqrt(diffsum) : std::sqrt(diffsum / lensum));
    if (dist > 1e-4f) {
      double ll = 0.0;
      int st = 0;
      getLogLikeBatch(par.data(), 1, par.getNParams(), &ll, &st, bonus_);
      meanParams = par;
    }
    par[nmp + 2] = static_cast<float>(bonus_[1]);
    par[nmp + 3] = static_cast<float>(bonus_[2]);
    par[nmp + 4] = static_cast<float>(bonus_[3]);
    par[nmp + 5] = static_cast<float>(bonus_[4]);
  }

  /* many walkers per launch: theta rows [knots..., sigmas..., offsets...,
     sigmult1, sigmult2, ...] */
  void getLogLikeBatch(const float* theta, unsigned int nwalk, unsigned int stride,
                       double* loglike, int* status, double* bonus = nullptr) const {
    if (sets.empty()) {
      for (unsigned w = 0; w < nwalk; ++w) loglike[w] = NAN;
      return;
    }
    const unsigned nmp = model.getNParams();
    if (stride < nmp + 2) throw affineExcept("calcLikeDouble", "getLogLike", "Not enough elements in params");
    if (!prepared) throw affineExcept("calcLikeDoubleSingle", "getLogLike", "R range not set");
    // the priors (src/calcLikeDouble.cc:1230-1280) are added on the device
    // after the dataset sum, for walkers whose datasets did not fail
    int rc = pushPriors();
    if (!rc)
      rc = pofd2d_loglike_batch(ctx.get(), theta, nwalk, stride, loglike,
                                reinterpret_cast<int32_t*>(status));
    if (rc) pofd_amd_detail::throw_status2(ctx.get(), rc, "calcLikeDouble", "getLogLike");
    if (bonus) {
      rc = pofd2d_get_bonus(ctx.get(), nwalk, bonus);
      if (rc) pofd_amd_detail::throw_status2(ctx.get(), rc, "calcLikeDouble", "getLogLike");
    }
  }

  /* the prior settings onto the context (pofd2d_set_priors) */
  int pushPriors() const {
    if (!ctx) return POFD_E_ARG;
    int flags = 0;
    if (has_sigma_prior1) flags |= POFD_PRIOR2_SIGMA1;
    if (has_sigma_prior2) flags |= POFD_PRIOR2_SIGMA2;
    if (has_cfirb[0]) flags |= POFD_PRIOR2_CFIRB1;
    if (has_cfirb[1]) flags |= POFD_PRIOR2_CFIRB2;
    if (has_poisson[0]) flags |= POFD_PRIOR2_POISSON1;
    if (has_poisson[1]) flags |= POFD_PRIOR2_POISSON2;
    const double sw[2] = {sigma_prior_width1, sigma_prior_width2};
    const double cf[4] = {cfirb_mean[0], cfirb_sigma[0], cfirb_mean[1], cfirb_sigma[1]};
    const double po[4] = {poisson_mean[0], poisson_sigma[0], poisson_mean[1], poisson_sigma[1]};
    return pofd2d_set_priors(ctx.get(), flags, sw, cf, po, regularization_alpha);
  }

  /* the device context (after setRRanges), for the sampler: its batches
     include the priors set here */
  pofd_ctx* context() const {
    if (!prepared) throw affineExcept("calcLikeDoubleSingle", "getLogLike", "R range not set");
    const int rc = pushPriors();
    if (rc) pofd_amd_detail::throw_status2(ctx.get(), rc, "calcLikeDouble", "context");
    return ctx.get();
  }

};

#endif
