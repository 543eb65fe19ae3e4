// Host-side, theta-independent preparation of beams and data.
//
// These run once per fit on the CPU, exactly as in the reference, and must
// be bit-exact with it (same glibc log2, same operation order), because the
// GPU kernels consume their outputs as-is:
//   beam::readFile pixel split/sort/totals/inverse   src/beam.cc:383-456
//   beam::makeHistogram                              src/beam.cc:467-556
//   fitsData::getMinMax/applyBinning/meanSubtract    src/fitsData.cc:348-437
#include "internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace pofd {

int Beam::set(const double* pixarr, size_t n, double pixsize_in,
              double MINVAL, std::string& err) {
  *this = Beam();
  pixsize = pixsize_in;
  minval = (MINVAL <= 0.0) ? 0.0 : MINVAL;          // src/beam.cc:307-310
  const double negminval = -minval;
  for (size_t i = 0; i < n; ++i) {                   // src/beam.cc:383-388
    if (std::isnan(pixarr[i]) || std::isinf(pixarr[i])) continue;
    if (pixarr[i] > minval) pos.push_back(pixarr[i]);
    if (pixarr[i] < negminval) neg.push_back(std::fabs(pixarr[i]));
  }
  if (pos.empty() && neg.empty()) {
    err = "beam::readFile: Beam file has only zero values";
    return POFD_E_ARG;
  }
  std::sort(pos.begin(), pos.end());                 // src/beam.cc:419-421
  std::sort(neg.begin(), neg.end());
  if (!pos.empty()) {                                // src/beam.cc:423-444
    totpos = pos[0];
    totpossq = totpos * totpos;
    for (size_t i = 1; i < pos.size(); ++i) {
      const double v = pos[i];
      totpos += v;
      totpossq += v * v;
    }
  }
  if (!neg.empty()) {
    totneg = neg[0];
    totnegsq = totneg * totneg;
    for (size_t i = 1; i < neg.size(); ++i) {
      const double v = neg[i];
      totneg += v;
      totnegsq += v * v;
    }
  }
  posinv.resize(pos.size());                         // src/beam.cc:446-456
  for (size_t i = 0; i < pos.size(); ++i) posinv[i] = 1.0 / pos[i];
  neginv.resize(neg.size());
  for (size_t i = 0; i < neg.size(); ++i) neginv[i] = 1.0 / neg[i];
  return POFD_OK;
}

// One sign of the log2-space histogram of the beam, src/beam.cc:493-524:
// bins in log2(pixel) padded by log2(1.001) on both sides; the stored value
// of a non-empty bin is the inverse of the mean pixel value in it.
static void histogram_one(const std::vector<double>& pix, uint32_t nbins,
                          std::vector<double>& weights,
                          std::vector<double>& invvals) {
  const double log2outscale = 0.0014419741739063218;  // log2(1.001)
  std::vector<uint32_t> wt(nbins, 0u);
  std::vector<double> val(nbins, 0.0);
  const double lo = log2(pix.front()) - log2outscale;
  const double hi = log2(pix.back()) + log2outscale;
  const double step = (hi - lo) / static_cast<double>(nbins);
  for (double v : pix) {
    const uint32_t idx = static_cast<uint32_t>((log2(v) - lo) / step);
    wt[idx] += 1;
    val[idx] += v;
  }
  weights.clear();
  invvals.clear();
  for (uint32_t i = 0; i < nbins; ++i)
    if (wt[i] > 0) {
      const double w = static_cast<double>(wt[i]);
      weights.push_back(w);
      invvals.push_back(val[i] / w);
    }
  for (double& v : invvals) v = 1.0 / v;
}

int Beam::make_histogram(uint32_t NBINS, std::string& err) {
  if (NBINS == 0) {
    err = "beam::makeHistogram: Invalid (non-positive) nbins";
    return POFD_E_ARG;
  }
  nbins = NBINS;
  is_pos_hist = is_neg_hist = false;
  posw.clear(); posv.clear(); negw.clear(); negv.clear();
  const uint32_t histothresh = 20;  // include/beam.h:31
  if (!pos.empty() && pos.size() >= histothresh) {
    histogram_one(pos, nbins, posw, posv);
    is_pos_hist = true;
  }
  if (!neg.empty() && neg.size() >= histothresh) {
    histogram_one(neg, nbins, negw, negv);
    is_neg_hist = true;
  }
  return POFD_OK;
}

void data_minmax(const double* d, size_t n, double& mn, double& mx) {
  mn = mx = d[0];
  for (size_t i = 1; i < n; ++i) {
    const double c = d[i];
    if (c < mn) mn = c;
    if (c > mx) mx = c;
  }
}

}  // namespace pofd

using pofd::Beam;

extern "C" int pofd_beam_create(pofd_beam** out, const double* pix,
                                uint64_t npix, double pixsize,
                                double minval) {
  if (!out || !pix || npix == 0) return POFD_E_ARG;
  auto* b = new (std::nothrow) pofd_beam();
  if (!b) return POFD_E_INTERNAL;
  std::string err;
  int rc = b->beam.set(pix, static_cast<size_t>(npix), pixsize, minval, err);
  if (rc) {
    delete b;
    *out = nullptr;
    return rc;
  }
  *out = b;
  return POFD_OK;
}

extern "C" int pofd_beam_make_histogram(pofd_beam* bm, uint32_t nbins) {
  if (!bm) return POFD_E_ARG;
  std::string err;
  return bm->beam.make_histogram(nbins, err);
}

extern "C" void pofd_beam_destroy(pofd_beam* bm) { delete bm; }

extern "C" int pofd_beam_get(const pofd_beam* bm, int which, double* out,
                             uint32_t* n) {
  if (!bm || !n) return POFD_E_ARG;
  const Beam& b = bm->beam;
  const std::vector<double>* v = nullptr;
  switch (which) {
    case 0: v = &b.pos; break;
    case 1: v = &b.neg; break;
    case 2: v = &b.posinv; break;
    case 3: v = &b.neginv; break;
    case 4: v = &b.posw; break;
    case 5: v = &b.posv; break;
    case 6: v = &b.negw; break;
    case 7: v = &b.negv; break;
    default: return POFD_E_ARG;
  }
  *n = static_cast<uint32_t>(v->size());
  if (out && !v->empty()) std::memcpy(out, v->data(), v->size() * sizeof(double));
  return POFD_OK;
}

extern "C" int pofd_beam_info(const pofd_beam* bm, double* info) {
  if (!bm || !info) return POFD_E_ARG;
  const Beam& b = bm->beam;
  const double conv = b.pixsize / 3600.0;
  for (int i = 0; i < 16; ++i) info[i] = 0.0;
  info[0] = static_cast<double>(b.pos.size());
  info[1] = static_cast<double>(b.neg.size());
  info[2] = b.pos.empty() ? 0.0 : 1.0;
  info[3] = b.neg.empty() ? 0.0 : 1.0;
  info[4] = b.pixsize;
  info[5] = b.minval;
  info[6] = b.is_pos_hist ? 1.0 : 0.0;
  info[7] = b.is_neg_hist ? 1.0 : 0.0;
  info[8] = static_cast<double>(b.posw.size());
  info[9] = static_cast<double>(b.negw.size());
  // beam::getEffectiveArea* src/beam.cc:614-668
  double area = 0.0, areapix = 0.0, areasq = 0.0;
  if (!b.pos.empty()) {
    area = b.totpos;
    areapix = b.totpos;
    areasq = b.totpossq * conv * conv;
  }
  if (!b.neg.empty()) {
    area += b.totneg;
    areapix += b.totneg;
    areasq += b.totnegsq * conv * conv;
  }
  info[10] = area * conv * conv;
  info[11] = b.pos.empty() ? 0.0 : b.totpos * conv * conv;
  info[12] = b.neg.empty() ? 0.0 : b.totneg * conv * conv;
  info[13] = areapix;
  info[14] = areasq;
  info[15] = b.pos.empty() ? NAN : b.pos.back();
  return POFD_OK;
}

extern "C" int pofd_data_bin(const double* d, uint64_t n, uint32_t nbins,
                             uint32_t* bins, double* bincent0,
                             double* bindelta) {
  // fitsData::applyBinning src/fitsData.cc:406-437
  if (!d || !bins || !bincent0 || !bindelta) return POFD_E_ARG;
  if (n == 0) return POFD_OK;
  if (nbins == 0) return POFD_E_ARG;
  std::memset(bins, 0, nbins * sizeof(uint32_t));
  double mn, mx;
  pofd::data_minmax(d, static_cast<size_t>(n), mn, mx);
  const double c0 = mn;
  const double delta = (nbins == 1) ? 2 * (mx - mn)
                                    : (mx - mn) / static_cast<double>(nbins - 1);
  const double ib = 1.0 / delta;
  for (uint64_t i = 0; i < n; ++i) {
    const uint32_t idx = static_cast<uint32_t>((d[i] - c0) * ib + 0.5);
    bins[idx] += 1;
  }
  *bincent0 = c0;
  *bindelta = delta;
  return POFD_OK;
}

extern "C" int pofd_data_mean_subtract(double* d, uint64_t n,
                                       double* mean_out) {
  // fitsData::getMean + meanSubtract src/fitsData.cc:373-401
  if (!d || n == 0) return POFD_E_ARG;
  double m = d[0];
  for (uint64_t i = 1; i < n; ++i) m += d[i];
  m /= static_cast<double>(n);
  for (uint64_t i = 0; i < n; ++i) d[i] -= m;
  if (mean_out) *mean_out = m;
  return POFD_OK;
}
