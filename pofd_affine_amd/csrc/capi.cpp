// C ABI: context, theta-independent device tables, batched launches.
//
// The host part restates, once per fit, the theta-independent pieces of the
// reference's PDFactory setup so that the kernels only do per-walker work:
//   calcLikeSingle::setSigmaBase like_offset        src/calcLike.cc:283-308
//   calcLikeSingle::setRRange                       src/calcLike.cc:346-371
//   numberCountsKnots::getRRangeInternal            src/numberCountsKnots.cc:345-365
//   PDFactory::initRFlux                            src/PDFactory.cc:361-414
//   PDFactory::initRInterp node fluxes              src/PDFactory.cc:446-463
//   PDFactory::initR fill index range               src/PDFactory.cc:562-574
//   GSL cspline_init tridiagonal factors (the node grid is fixed)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "internal.h"
#include "internal2d.h"

namespace {

using pofd::BeamsetDev;
using pofd::FieldDev;
using pofd::WorkDev;

const double kPi = 3.141592653589793238462643383279502884197;

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

struct FieldHost {
  std::vector<double> pix;        // sorted ascending (unbinned)
  std::vector<uint32_t> bins;     // binned
  uint64_t npix = 0;
  bool binned = false;
  double bincent0 = 0, bindelta = 0;
  double sigma_base = 0, like_norm = 1, like_offset = 0;
  double mindata = 0, maxdata = 0;
  std::unique_ptr<DevBuf> dpix, dbins;
};

struct Beamset {
  pofd::Beam beam;
  std::vector<FieldHost> fields;
  double minRFlux = NAN, maxRFlux = NAN, maxsigma_base = NAN;
  double minDataFlux = NAN, maxDataFlux = NAN;
  bool prepared = false;
  BeamsetDev dev{};
  std::vector<std::unique_ptr<DevBuf>> tables;
  // workspace
  WorkDev wk{};
  std::unique_ptr<DevBuf> wbuf;
  uint32_t wcap = 0;
};

}  // namespace

struct pofd_ctx {
  int device = 0;
  std::string err;
  double spectral_cut = 64.0 * 0.69314718055994530942;
  bool configured = false;
  uint32_t nflux = 0, ninterp = 0;
  std::vector<double> knots, logknots;
  double exp_conf = 0.0;
  std::vector<std::unique_ptr<Beamset>> sets;
  uint32_t last_nwalk = 0;
  int force_parts = 0;  // workgroups per walker: 0 = by batch size
  // measured workgroups per walker by batch size (pofd1d_autotune_parts):
  // host copy and device copy, both empty until the first tuning
  std::array<uint8_t, pofd::kMaxSlots + 1> parts_tab{};
  std::unique_ptr<DevBuf> dparts;
  bool parts_tuned = false;
  pofd::PriorDev priors{};  // added on the device after the beam sets
  std::unique_ptr<DevBuf> dtheta, dll, dst;
  size_t cap_theta = 0, cap_out = 0;
  // stand-alone PD / PDDouble likelihoods: the P(D), the data and partials
  std::unique_ptr<DevBuf> pdl_pd, pdl_d1, pdl_d2, pdl_bins, pdl_part;
  // stage timing: one set of POFD_NSTAGES+1 events per beam-set launch
  bool stage_timing = false;
  std::vector<std::array<hipEvent_t, POFD_NSTAGES + 1>> stage_events;
  size_t stage_used = 0;
  pofd2::Ctx2* c2 = nullptr;  // two-band state (capi2d.cpp)
  // bumped whenever device tables or workspaces are replaced, or a launch
  // parameter captured by value changes; a captured graph records it and
  // refuses to launch once it differs
  uint64_t generation = 0;
  int noalloc_rc = POFD_E_INTERNAL;  // pofd::noalloc_init's result
  ~pofd_ctx() {
    for (auto& ev : stage_events)
      for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    pofd2::ctx2_free(c2);
  }
};

// accessors for the two-band C ABI (capi2d.cpp)
namespace pofd2 {
Ctx2*& ctx2_of(pofd_ctx* c) { return c->c2; }
int ctx_device(const pofd_ctx* c) { return c->device; }
double ctx_spectral_cut(const pofd_ctx* c) { return c->spectral_cut; }
int ctx_fail(pofd_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}
// the next set of POFD_NSTAGES+1 stage events while stage timing is on
// (nullptr when it is off or the tally is full)
int ctx_stage_events(pofd_ctx* c, hipEvent_t** ev) {
  constexpr size_t kMaxTimedLaunches = 4096;
  *ev = nullptr;
  if (!c->stage_timing || c->stage_used >= kMaxTimedLaunches) return POFD_OK;
  if (c->stage_used == c->stage_events.size()) {
    std::array<hipEvent_t, POFD_NSTAGES + 1> fresh{};
    for (auto& e : fresh) {
      hipError_t he = hipEventCreate(&e);
      if (he != hipSuccess)
        return ctx_fail(c, POFD_E_INTERNAL,
                        std::string("pofd::hip: hipEventCreate: ") + hipGetErrorString(he));
    }
    c->stage_events.push_back(fresh);
  }
  *ev = c->stage_events[c->stage_used++].data();
  return POFD_OK;
}
}  // namespace pofd2

namespace {

int fail(pofd_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_fail(pofd_ctx* c, hipError_t e, const char* what) {
  return fail(c, POFD_E_INTERNAL,
              std::string("pofd::hip: ") + what + ": " + hipGetErrorString(e));
}

int dev_alloc(pofd_ctx* c, std::unique_ptr<DevBuf>& b, size_t bytes) {
  if (b && b->bytes >= bytes) return POFD_OK;
  b.reset(new DevBuf());
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(&b->p, bytes);
  if (e != hipSuccess) {
    b.reset();
    return hip_fail(c, e, "hipMalloc");
  }
  b->bytes = bytes;
  return POFD_OK;
}

template <typename T>
int upload(pofd_ctx* c, Beamset& s, const std::vector<T>& v, const T** out) {
  s.tables.emplace_back(new DevBuf());
  DevBuf& b = *s.tables.back();
  const size_t bytes = std::max<size_t>(v.size() * sizeof(T), 8);
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) return hip_fail(c, e, "hipMalloc(table)");
  b.bytes = bytes;
  if (!v.empty()) {
    e = hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(table)");
  }
  *out = static_cast<const T*>(b.p);
  return POFD_OK;
}

// gsl_interp_bsearch over [0, n-1] with the accelerator's clamping
size_t bsearch_idx(const double* x, size_t n, double v) {
  if (v <= x[0]) return 0;
  if (v >= x[n - 1]) return n - 2;
  size_t lo = 0, hi = n - 1;
  while (hi > lo + 1) {
    const size_t i = (hi + lo) / 2;
    if (x[i] > v)
      hi = i;
    else
      lo = i;
  }
  return lo;
}

}  // namespace

// Gauss-Legendre nodes and weights on [-1, 1] (Newton on the Legendre
// recurrence in long double)
void pofd::gauss_legendre(int n, double* x, double* w) {
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int i = 0; i < n; ++i) {
    long double z = cosl(pi * (i + 0.75L) / (n + 0.5L)), dp = 0.0L;
    for (int it = 0; it < 100; ++it) {
      long double p0 = 1.0L, p1 = z;
      for (int k = 2; k <= n; ++k) {
        const long double p2 = ((2.0L * k - 1.0L) * z * p1 - (k - 1.0L) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0L);
      const long double dz = p1 / dp;
      z -= dz;
      if (fabsl(dz) < 1e-19L) break;
    }
    x[i] = static_cast<double>(z);
    w[i] = static_cast<double>(2.0L / ((1.0L - z * z) * dp * dp));
  }
}

namespace {

std::vector<double2> twiddles(size_t n, size_t count) {
  std::vector<double2> w(count);
  for (size_t s = 0; s < count; ++s) {
    const long double a = -2.0L * static_cast<long double>(kPi) *
                          static_cast<long double>(s) / static_cast<long double>(n);
    w[s] = make_double2(static_cast<double>(cosl(a)), static_cast<double>(sinl(a)));
  }
  return w;
}

int prepare_beamset(pofd_ctx* c, Beamset& s) {
  const pofd::Beam& bm = s.beam;
  const uint32_t N = c->nflux, ninterp = c->ninterp;
  const unsigned nk = static_cast<unsigned>(c->knots.size());
  if (s.fields.empty())
    return fail(c, POFD_E_ARG, "calcLikeSingle::readDataFromFile: No data sets");
  const bool haspos = !bm.pos.empty(), hasneg = !bm.neg.empty();
  if (!haspos && !hasneg)
    return fail(c, POFD_E_ARG, "PDFactory::initPD: Empty beam");
  // data range (src/calcLike.cc:205-230) and sigma base (:283-308)
  s.minDataFlux = s.fields[0].mindata;
  s.maxDataFlux = s.fields[0].maxdata;
  s.maxsigma_base = s.fields[0].sigma_base;
  for (auto& f : s.fields) {
    s.minDataFlux = std::min(s.minDataFlux, f.mindata);
    s.maxDataFlux = std::max(s.maxDataFlux, f.maxdata);
    if (f.sigma_base > s.maxsigma_base) s.maxsigma_base = f.sigma_base;
  }
  for (auto& f : s.fields) {
    const double var = f.sigma_base * f.sigma_base + c->exp_conf * c->exp_conf;
    f.like_offset = 0.0;
    if (var > 0) {
      const double ndd = static_cast<double>(f.npix);
      f.like_offset = -0.5 * ndd * (log(2.0 * kPi * var) + 1.0);
    }
  }
  // range where R is non-zero (numberCountsKnots::getRRangeInternal,
  // src/numberCountsKnots.cc:345-365)
  const double minknot = c->knots[0], maxknot = c->knots[nk - 1];
  double rmin, rmax;
  if (haspos) {
    rmax = maxknot * bm.pos.back();
    rmin = hasneg ? -maxknot * bm.neg.back() : 0.0;
  } else {
    rmax = 0.0;
    rmin = -maxknot * bm.neg.back();
  }
  // setRRange (src/calcLike.cc:346-371)
  if (std::isnan(s.minRFlux) || std::isnan(s.maxRFlux)) {
    const double sigma = sqrt(c->exp_conf * c->exp_conf +
                              s.maxsigma_base * s.maxsigma_base);
    s.minRFlux = 1.05 * rmin;
    s.maxRFlux = 1.05 * rmax + 7.5 * sigma;
    if (s.minDataFlux < s.minRFlux && hasneg) s.minRFlux = s.minDataFlux;
    if (s.maxDataFlux > s.maxRFlux && haspos) s.maxRFlux = s.maxDataFlux;
  }
  // initRFlux (src/PDFactory.cc:361-414): negative fluxes wrap to the top
  double minflux = s.minRFlux, maxflux = s.maxRFlux;
  if (maxflux < minflux) std::swap(minflux, maxflux);
  if (!(maxflux > 0))
    return fail(c, POFD_E_ARG, "PDFactory::initPD: Invalid (non-positive) maxflux");
  if (minflux > 0) minflux = 0.0;
  const double inm1 = 1.0 / static_cast<double>(N - 1);
  int wrapRidx;
  double minflux_R, dflux;
  if (minflux == 0) {
    wrapRidx = static_cast<int>(N - 1);
    minflux_R = 0.0;
    dflux = maxflux * inm1;
  } else {
    dflux = (maxflux - minflux) * inm1;
    dflux = maxflux / (N - floor(-minflux / dflux) - 1.0);
    const double wr = maxflux / dflux + 0.9999999999;
    if (!(wr >= 1.0 && wr < static_cast<double>(N)))
      return fail(c, POFD_E_ARG, "PDFactory::initRFlux: R flux range too "
                                 "lopsided for the transform size");
    wrapRidx = static_cast<int>(static_cast<unsigned>(wr));
    minflux_R = (wrapRidx + 1) * dflux - static_cast<double>(N) * dflux;
  }
  // fluxes of grid points above wrapRidx: i * dflux + mval (initR)
  const double mval = minflux_R - static_cast<double>(wrapRidx + 1) * dflux;

  // initRInterp node fluxes (src/PDFactory.cc:426-527): segment 0 the
  // positive beam, segment 1 the negative beam (ascending fluxes < 0)
  const double ininterpm1 = 1.0 / static_cast<double>(ninterp - 1);
  const double subedgemult = 1e-5;  // include/PDFactory.h:29
  const int nseg = hasneg ? 2 : 1;
  std::vector<double> xnode(static_cast<size_t>(nseg) * ninterp, 0.0);
  if (haspos) {
    const double mininterpflux = minknot * subedgemult;
    double maxinterpflux = maxknot * bm.pos.back();
    const double dinterp = 0.001 * (maxinterpflux - mininterpflux) * ininterpm1;
    maxinterpflux -= dinterp;
    const double dinterpflux = (log2(maxinterpflux / mininterpflux)) * ininterpm1;
    for (uint32_t i = 0; i < ninterp; ++i)
      xnode[i] = mininterpflux * exp2(static_cast<double>(i) * dinterpflux);
  } else {  // unused segment, kept increasing for the tables below
    for (uint32_t i = 0; i < ninterp; ++i) xnode[i] = 1.0 + i;
  }
  if (hasneg) {
    double maxinterpflux = -minknot * subedgemult;
    double mininterpflux = -maxknot * bm.neg.back();
    const double dinterp = 0.001 * (maxinterpflux - mininterpflux) * ininterpm1;
    mininterpflux += dinterp;
    maxinterpflux -= dinterp;
    const double dinterpflux = (log2(mininterpflux / maxinterpflux)) * ininterpm1;
    for (uint32_t i = 0; i < ninterp; ++i)
      xnode[ninterp + i] =
          maxinterpflux * exp2(static_cast<double>(ninterp - i - 1) * dinterpflux);
  }
  for (int g = 0; g < nseg; ++g)
    for (uint32_t i = 1; i < ninterp; ++i)
      if (!(xnode[g * ninterp + i] > xnode[g * ninterp + i - 1]))
        return fail(c, POFD_E_ARG, "PDFactory::initRInterp: interpolation "
                                   "fluxes not increasing");

  // beam entries per segment: histogram (weights, inverse bin means) or raw
  // inverse pixel values with weight 1
  std::vector<double> bw, bieta;
  auto add_entries = [&](bool ishist, const std::vector<double>& w,
                         const std::vector<double>& v, const std::vector<double>& inv) {
    if (ishist) {
      bw.insert(bw.end(), w.begin(), w.end());
      bieta.insert(bieta.end(), v.begin(), v.end());
    } else {
      bw.insert(bw.end(), inv.size(), 1.0);
      bieta.insert(bieta.end(), inv.begin(), inv.end());
    }
  };
  if (haspos) add_entries(bm.is_pos_hist, bm.posw, bm.posv, bm.posinv);
  const size_t nbeam = bw.size();
  if (hasneg) add_entries(bm.is_neg_hist, bm.negw, bm.negv, bm.neginv);
  const size_t nbeamn = bw.size() - nbeam;
  // knot-interval table: for every (beam entry, node) the knot interval of
  // log2(|flux| * inverse beam) and its offset, or -1 where getR skips it
  // (array getR, src/numberCountsKnotsSpline.cc:436-529)
  std::vector<int8_t> rk(bw.size() * ninterp, -1);
  std::vector<double> rdelx(bw.size() * ninterp, 0.0);
  for (size_t j = 0; j < bw.size(); ++j) {
    const bool negseg = j >= nbeam;
    const double* xs = xnode.data() + (negseg ? ninterp : 0);
    for (uint32_t i = 0; i < ninterp; ++i) {
      const double fd = xs[i];
      if (negseg ? !(fd < 0 && fd >= rmin) : !(fd > 0 && fd <= rmax)) continue;
      const double currarg = (negseg ? -fd : fd) * bieta[j];
      if (currarg < minknot || currarg >= maxknot) continue;
      const double lx = log2(currarg);
      const size_t k = bsearch_idx(c->logknots.data(), nk, lx);
      rk[j * ninterp + i] = static_cast<int8_t>(k);
      rdelx[j * ninterp + i] = lx - c->logknots[k];
    }
  }
  // fixed LDL^T factors of each segment's node spline system (GSL
  // cspline_init + gsl_linalg_solve_symm_tridiag)
  const uint32_t sys = ninterp - 2;
  std::vector<double> ginv(nseg * (ninterp - 1)), lalpha(nseg * sys), lgamma(nseg * sys, 0.0);
  int kwarm[2] = {pofd::kWarm, pofd::kWarm};
  for (int g = 0; g < nseg; ++g) {
    const double* xs = xnode.data() + g * ninterp;
    double* gi = ginv.data() + g * (ninterp - 1);
    double* la = lalpha.data() + g * sys;
    double* lg = lgamma.data() + g * sys;
    std::vector<double> diag(sys), off(sys);
    for (uint32_t i = 0; i + 1 < ninterp; ++i) {
      const double h = xs[i + 1] - xs[i];
      gi[i] = (h != 0.0) ? 1.0 / h : 0.0;
    }
    for (uint32_t i = 0; i < sys; ++i) {
      const double h_i = xs[i + 1] - xs[i];
      const double h_ip1 = xs[i + 2] - xs[i + 1];
      off[i] = h_ip1;
      diag[i] = 2.0 * (h_ip1 + h_i);
    }
    la[0] = diag[0];
    lg[0] = off[0] / la[0];
    for (uint32_t i = 1; i + 1 < sys; ++i) {
      la[i] = diag[i] - off[i - 1] * lg[i - 1];
      lg[i] = off[i] / la[i];
    }
    if (sys > 1) la[sys - 1] = diag[sys - 1] - off[sys - 2] * lg[sys - 2];
    // warm-up of the kernel's truncated recurrences: a start k steps early
    // is off by at most max(lg)^k of the skipped value, 2^-64 at this k
    // (uniform nodes: max(lg) = 2 - sqrt(3) = 0.268, 34 steps)
    double lgmax = 0.0;
    for (uint32_t i = 0; i < sys; ++i) lgmax = std::max(lgmax, std::fabs(lg[i]));
    int kw = pofd::kWarm;
    if (lgmax < 0.5) {
      kw = (lgmax > 0.0) ? static_cast<int>(std::ceil(64.0 / -std::log2(lgmax))) : 1;
      kw = std::min(std::max(kw, 1), pofd::kWarm);
    }
    kwarm[g] = kw;
  }
  // fixed factors of the knot spline's system (the knot positions do not
  // change with theta): h_i, 1/h_i, and the LDL^T diagonal and multipliers
  // of GSL's cspline_init + gsl_linalg_solve_symm_tridiag, in its order
  std::vector<double> kfac(4 * pofd::kMaxKnots, 0.0);
  {
    const int nkk = static_cast<int>(nk), ksys = nkk - 2;
    const double* xa = c->logknots.data();
    double* kh = kfac.data();
    double* kgi = kh + pofd::kMaxKnots;
    double* kal = kgi + pofd::kMaxKnots;
    double* kgm = kal + pofd::kMaxKnots;
    for (int i = 0; i + 1 < nkk; ++i) {
      kh[i] = xa[i + 1] - xa[i];
      kgi[i] = (kh[i] != 0.0) ? 1.0 / kh[i] : 0.0;
    }
    if (ksys >= 1) {
      kal[0] = 2.0 * (kh[1] + kh[0]);
      if (ksys > 1) {
        kgm[0] = kh[1] / kal[0];
        for (int i = 1; i < ksys - 1; ++i) {
          kal[i] = 2.0 * (kh[i + 1] + kh[i]) - kh[i] * kgm[i - 1];
          kgm[i] = kh[i + 1] / kal[i];
        }
        kal[ksys - 1] = 2.0 * (kh[ksys] + kh[ksys - 1]) - kh[ksys - 1] * kgm[ksys - 2];
      }
    }
  }
  // initR fill ranges (src/PDFactory.cc:560-613); empty ranges are [1, 0]
  const bool fillpos = haspos && (maxflux > 0);
  const bool fillneg = hasneg && (minflux < 0) && (wrapRidx > 0);
  long minit = 1, maxit = 0, minitn = 1, maxitn = 0;
  if (fillpos) {
    const double minI = xnode[0], maxI = xnode[ninterp - 1];
    minit = static_cast<long>(static_cast<unsigned>(minI / dflux + 0.9999999999999999));
    maxit = static_cast<long>(static_cast<unsigned>(maxI / dflux));
    if (maxit > wrapRidx) maxit = wrapRidx;
  }
  if (fillneg) {
    const double minI = xnode[ninterp], maxI = xnode[2 * ninterp - 1];
    const double lo = (minI - minflux_R) / dflux + wrapRidx + 1 + 0.9999999999999999;
    const double hi = (maxI - minflux_R) / dflux + wrapRidx + 1;
    minitn = (lo > 0) ? static_cast<long>(static_cast<unsigned>(lo)) : 0;
    if (minitn <= wrapRidx) minitn = wrapRidx + 1;
    maxitn = (hi > 0) ? static_cast<long>(static_cast<unsigned>(hi)) : -1;
    if (maxitn > static_cast<long>(N) - 1) maxitn = N - 1;
  }
  // node interval of every flux of the grid (what gsl_interp_accel_find
  // returns for it in initR), as segment * ninterp + interval
  std::vector<uint16_t> kst(N, 0);
  for (uint32_t b = 0; b < N; ++b) {
    if (static_cast<int>(b) <= wrapRidx) {
      if (!fillpos) continue;
      const double cf = static_cast<double>(b) * dflux;
      kst[b] = static_cast<uint16_t>(bsearch_idx(xnode.data(), ninterp, cf));
    } else {
      if (!fillneg) continue;
      const double cf = static_cast<double>(b) * dflux + mval;
      kst[b] = static_cast<uint16_t>(ninterp + bsearch_idx(xnode.data() + ninterp, ninterp, cf));
    }
  }
  const uint32_t M = N / 2, M2 = M / pofd::kM1;

  BeamsetDev& d = s.dev;
  std::memset(&d, 0, sizeof(d));
  d.N = static_cast<int>(N);
  d.M = static_cast<int>(M);
  d.M2 = static_cast<int>(M2);
  d.ninterp = static_cast<int>(ninterp);
  d.nknots = static_cast<int>(nk);
  d.nbeam = static_cast<int>(nbeam);
  d.nbeamn = static_cast<int>(nbeamn);
  d.nseg = nseg;
  d.kwarm[0] = kwarm[0];
  d.kwarm[1] = kwarm[1];
  d.haspos = haspos ? 1 : 0;
  d.minit = static_cast<int>(minit);
  d.maxit = static_cast<int>(maxit);
  d.minitn = static_cast<int>(minitn);
  d.maxitn = static_cast<int>(maxitn);
  d.wrapRidx = wrapRidx;
  d.dflux = dflux;
  d.minflux_R = minflux_R;
  d.mval = mval;
  d.logfac = log2(10.0);
  double prefac = bm.pixsize / 3600.0;
  d.prefac = prefac * prefac;
  d.maxsigma_base = s.maxsigma_base;
  d.iflux = (2.0 * kPi) / (N * dflux);
  d.spectral_cut = c->spectral_cut;
  for (unsigned i = 0; i < nk; ++i) d.logknots[i] = c->logknots[i];
  pofd::gauss_legendre(pofd::kGL, d.glx, d.glw);
  s.tables.clear();
  int rc;
  if ((rc = upload(c, s, xnode, &d.xnode))) return rc;
  if ((rc = upload(c, s, ginv, &d.ginv))) return rc;
  if ((rc = upload(c, s, lalpha, &d.lalpha))) return rc;
  if ((rc = upload(c, s, lgamma, &d.lgamma))) return rc;
  if ((rc = upload(c, s, kfac, &d.kfac))) return rc;
  if ((rc = upload(c, s, rk, &d.rk))) return rc;
  if ((rc = upload(c, s, rdelx, &d.rdelx))) return rc;
  if ((rc = upload(c, s, bw, &d.bw))) return rc;
  if ((rc = upload(c, s, bieta, &d.bieta))) return rc;
  if ((rc = upload(c, s, kst, &d.kst))) return rc;
  const std::vector<double2> w256 = twiddles(256, 256), wlo = twiddles(M, 256),
                             wm2 = twiddles(M2, M2);
  if ((rc = upload(c, s, w256, &d.W256))) return rc;
  if ((rc = upload(c, s, wlo, &d.Wlo))) return rc;
  if ((rc = upload(c, s, wm2, &d.Wm2))) return rc;
  // W_N^t - 1 for t in [0, N/2] (the rest by conjugate symmetry), formed
  // without cancellation as (-2 sin^2(pi t/N), -sin(2 pi t/N)) in long
  // double: the direct sums over the head of the grid use it
  {
    std::vector<double2> em1(N / 2 + 1);
    for (size_t t = 0; t <= N / 2; ++t) {
      const long double a = static_cast<long double>(kPi) * static_cast<long double>(t) /
                            static_cast<long double>(N);
      const long double sa = sinl(a);
      em1[t] = make_double2(static_cast<double>(-2.0L * sa * sa),
                            static_cast<double>(-sinl(2.0L * a)));
    }
    if ((rc = upload(c, s, em1, &d.em1))) return rc;
  }
  // the head of the grid: the first half transform row (and the last half
  // row, which holds the wrapped fluxes just below zero, when there are any)
  d.headn = static_cast<int>(M2);
  d.headneg = (minitn <= maxitn && maxitn >= static_cast<long>(N) - d.headn) ? 1 : 0;
  d.nfields = static_cast<int>(s.fields.size());
  for (size_t f = 0; f < s.fields.size(); ++f) {
    FieldHost& fh = s.fields[f];
    FieldDev& fd = d.fields[f];
    fd.sigma_base = fh.sigma_base;
    fd.like_offset = fh.like_offset;
    fd.like_norm = fh.like_norm;
    fd.binned = fh.binned ? 1 : 0;
    fd.nbins = static_cast<uint32_t>(fh.bins.size());
    fd.bincent0 = fh.bincent0;
    fd.bindelta = fh.bindelta;
    fd.npix = fh.binned ? 0 : fh.pix.size();
    if (fh.binned) {
      if ((rc = dev_alloc(c, fh.dbins, fh.bins.size() * sizeof(uint32_t)))) return rc;
      hipError_t e = hipMemcpy(fh.dbins->p, fh.bins.data(),
                               fh.bins.size() * sizeof(uint32_t),
                               hipMemcpyHostToDevice);
      if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(bins)");
      fd.bins = static_cast<const uint32_t*>(fh.dbins->p);
      fd.pix = nullptr;
    } else {
      if ((rc = dev_alloc(c, fh.dpix, fh.pix.size() * sizeof(double)))) return rc;
      hipError_t e = hipMemcpy(fh.dpix->p, fh.pix.data(),
                               fh.pix.size() * sizeof(double),
                               hipMemcpyHostToDevice);
      if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(pix)");
      fd.pix = static_cast<const double*>(fh.dpix->p);
      fd.bins = nullptr;
    }
  }
  s.wcap = 0;  // workspace layout depends on the tables
  s.prepared = true;
  ++c->generation;
  return POFD_OK;
}

// Per-walker workspace of a beam set for batches of up to nwalk walkers:
// the strides and pointers into `base` (nullptr: sizes only).  Returns the
// number of doubles it spans.
size_t workspace_layout(const BeamsetDev& d, uint32_t nwalk, double* base, WorkDev& w) {
  const size_t nf = static_cast<size_t>(d.nfields);
  auto r8 = [](size_t n) { return (n + 7) & ~size_t(7); };  // 64-byte units
  w.coef_stride = r8(static_cast<size_t>(d.nseg) * d.ninterp * 4);
  w.ynode_stride = r8(static_cast<size_t>(d.nseg) * d.ninterp);
  // forward rows, then the noisy rows of each dataset
  w.zrows_stride = r8((1 + nf) * static_cast<size_t>(d.M) * 2) / 2;
  w.spec_stride = r8(2 * (static_cast<size_t>(d.M) + 1)) / 2;
  w.pofd_stride = r8(static_cast<size_t>(d.N));
  w.blk_stride = r8(static_cast<size_t>(d.N / pofd::kBlk));
  // (walker, part) slots for batches split over several workgroups per
  // walker: at most 4 parts, and walkers x parts <= 256
  const uint32_t nslots = std::min<uint32_t>(pofd::kMaxSlots, 4 * nwalk);
  w.zpart_stride = r8(2 * static_cast<size_t>(d.M)) / 2;
  const size_t n_coef = w.coef_stride * nwalk;
  const size_t n_ynode = w.ynode_stride * nwalk;
  const size_t n_z = 2 * w.zrows_stride * nwalk;
  const size_t n_spec = 2 * w.spec_stride * nwalk;  // prtrans
  const size_t n_pofd = w.pofd_stride * nf * nwalk;
  const size_t n_blk = 2 * w.blk_stride * nf * nwalk;
  const size_t n_dbg = r8(pofd::kDbg) * nf * nwalk;
  const size_t n_state = r8(pofd::kState) * nwalk;
  const size_t n_upart = r8(pofd::kPixBlocks) * nf * nwalk;
  const size_t n_bonus = r8(4 * static_cast<size_t>(nwalk));
  const size_t n_zpart = 2 * w.zpart_stride * nslots;
  const size_t n_mpart = r8(2 * static_cast<size_t>(nslots));
  const size_t n_ipart = r8(static_cast<size_t>(nslots) * nf * pofd::kIPart);
  const size_t total = n_coef + n_ynode + n_z + n_spec + n_pofd + n_blk + n_dbg + n_state +
                       n_upart + n_bonus + n_zpart + n_mpart + n_ipart;
  w.nslots = nslots;
  if (!base) return total;
  w.coef = base;
  base += n_coef;
  w.ynode = base;
  base += n_ynode;
  w.zrows = reinterpret_cast<double2*>(base);
  base += n_z;
  w.prtrans = reinterpret_cast<double2*>(base);
  base += n_spec;
  w.pofd = base;
  base += n_pofd;
  w.bmin = base;
  base += n_blk / 2;
  w.bsum = base;
  base += n_blk / 2;
  w.dbg = base;
  base += n_dbg;
  w.state = base;
  base += n_state;
  w.upart = base;
  base += n_upart;
  w.bonus = base;
  base += n_bonus;
  w.zpart = reinterpret_cast<double2*>(base);
  base += n_zpart;
  w.mpart = base;
  base += n_mpart;
  w.ipart = base;
  return total;
}

int ensure_workspace(pofd_ctx* c, Beamset& s, uint32_t nwalk) {
  if (s.wcap >= nwalk && s.wbuf) return POFD_OK;
  WorkDev w{};
  const size_t total = workspace_layout(s.dev, nwalk, nullptr, w);
  s.wbuf.reset();
  int rc = dev_alloc(c, s.wbuf, total * sizeof(double));
  if (rc) return rc;
  workspace_layout(s.dev, nwalk, static_cast<double*>(s.wbuf->p), s.wk);
  s.wcap = nwalk;
  ++c->generation;
  return POFD_OK;
}


int check_ready(pofd_ctx* c) {
  if (!c) return POFD_E_ARG;
  if (!c->configured)
    return fail(c, POFD_E_ARG, "calcLike::getLogLike: pofd1d_configure not called");
  if (c->sets.empty())
    return fail(c, POFD_E_ARG, "calcLike::getLogLike: no beam sets");
  for (auto& s : c->sets)
    if (!s->prepared)
      return fail(c, POFD_E_ARG,
                  "calcLikeSingle::getNegLogLike: R range not set "
                  "(pofd1d_prepare not called)");
  return POFD_OK;
}

}  // namespace

extern "C" int pofd_create(pofd_ctx** out, int hip_device) {
  if (!out) return POFD_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= hip_device ||
      hip_device < 0)
    return POFD_E_NOGPU;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, hip_device) != hipSuccess)
    return POFD_E_NOGPU;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return POFD_E_NOGPU;
  if (hipSetDevice(hip_device) != hipSuccess) return POFD_E_NOGPU;
  auto* c = new (std::nothrow) pofd_ctx();
  if (!c) return POFD_E_INTERNAL;
  c->device = hip_device;
  // the allocation-free entry point's scratch on this device (its address
  // is resolved once per device); a failure is reported by that entry point
  c->noalloc_rc = pofd::noalloc_init(hip_device);
  *out = c;
  return POFD_OK;
}

extern "C" void pofd_destroy(pofd_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipDeviceSynchronize();
  delete c;
}

extern "C" const char* pofd_last_error(const pofd_ctx* c) {
  return c ? c->err.c_str() : "null context";
}

extern "C" int pofd_set_option(pofd_ctx* c, int option, double value) {
  if (!c) return POFD_E_ARG;
  if (option == POFD_OPT_SPECTRAL_CUT) {
    if (!(value > 0.0))
      return fail(c, POFD_E_ARG, "pofd_set_option: spectral cut must be > 0");
    c->spectral_cut = value;
    for (auto& s : c->sets) s->dev.spectral_cut = value;
    ++c->generation;
    return POFD_OK;
  }
  if (option == POFD_OPT_STAGE_TIMING) {
    c->stage_timing = value != 0.0;
    c->stage_used = 0;
    return POFD_OK;
  }
  return fail(c, POFD_E_ARG, "pofd_set_option: unknown option");
}

extern "C" int pofd1d_configure(pofd_ctx* c, uint32_t nflux, uint32_t ninterp,
                                const double* knot_pos, uint32_t nknots,
                                double exp_conf) {
  if (!c) return POFD_E_ARG;
  if (nflux < 8192 || nflux > 262144 || (nflux & (nflux - 1)) != 0)
    return fail(c, POFD_E_ARG,
                "pofd1d_configure: nflux must be a power of two from 2^13 to 2^18");
  if (ninterp < 16 || ninterp > 2048)
    return fail(c, POFD_E_ARG, "pofd1d_configure: ninterp must be in [16, 2048]");
  if (!knot_pos || nknots < 2 || nknots > static_cast<uint32_t>(pofd::kMaxKnots))
    return fail(c, POFD_E_ARG, "pofd1d_configure: 2 to 64 knots required");
  for (uint32_t i = 0; i < nknots; ++i)
    if (!(knot_pos[i] > 0.0))
      return fail(c, POFD_E_ARG,
                  "numberCountsKnots::setKnots: Negative knot positions not "
                  "allowed");
  for (uint32_t i = 1; i < nknots; ++i)
    if (!(knot_pos[i] > knot_pos[i - 1]))
      return fail(c, POFD_E_ARG,
                  "numberCountsKnots::setKnotPositions: Knot positions not "
                  "monotonically increasing");
  c->nflux = nflux;
  c->ninterp = ninterp;
  c->knots.assign(knot_pos, knot_pos + nknots);
  c->logknots.resize(nknots);
  for (uint32_t i = 0; i < nknots; ++i) c->logknots[i] = log2(c->knots[i]);
  c->exp_conf = exp_conf;
  c->configured = true;
  for (auto& s : c->sets) s->prepared = false;
  ++c->generation;
  return POFD_OK;
}

extern "C" int pofd1d_add_beamset(pofd_ctx* c, const pofd_beam* bm, int* id) {
  if (!c || !bm || !id) return POFD_E_ARG;
  std::unique_ptr<Beamset> s(new Beamset());
  s->beam = bm->beam;
  c->sets.push_back(std::move(s));
  *id = static_cast<int>(c->sets.size() - 1);
  return POFD_OK;
}

extern "C" int pofd1d_add_field(pofd_ctx* c, int b, const double* pix,
                                uint64_t npix, const uint32_t* bins,
                                uint32_t nbins, double bincent0,
                                double bindelta, double sigma_base,
                                double like_norm, int* fid) {
  if (!c || b < 0 || static_cast<size_t>(b) >= c->sets.size())
    return fail(c, POFD_E_ARG, "pofd1d_add_field: bad beam set");
  Beamset& s = *c->sets[b];
  if (s.fields.size() >= static_cast<size_t>(pofd::kMaxFields))
    return fail(c, POFD_E_ARG, "pofd1d_add_field: too many fields");
  if (npix == 0)
    return fail(c, POFD_E_ARG,
                "calcLikeSingle::readDataFromFile: No unmasked pixels in data");
  if (!(like_norm != 0.0))
    return fail(c, POFD_E_ARG, "pofd1d_add_field: like_norm must be non-zero");
  FieldHost f;
  f.sigma_base = sigma_base;
  f.like_norm = like_norm;
  f.npix = npix;
  if (bins) {
    if (nbins == 0) return fail(c, POFD_E_ARG, "fitsData::applyBinning: no bins");
    f.binned = true;
    f.bins.assign(bins, bins + nbins);
    f.bincent0 = bincent0;
    f.bindelta = bindelta;
    // fitsData::getMinMax binned branch src/fitsData.cc:353-357
    f.mindata = bincent0;
    f.maxdata = bincent0 + static_cast<double>(nbins - 1) * bindelta;
  } else {
    if (!pix) return fail(c, POFD_E_ARG, "pofd1d_add_field: no pixels");
    for (uint64_t i = 0; i < npix; ++i)
      if (!std::isfinite(pix[i]))
        return fail(c, POFD_E_ARG, "pofd1d_add_field: non-finite pixel value");
    f.pix.assign(pix, pix + npix);
    pofd::data_minmax(f.pix.data(), f.pix.size(), f.mindata, f.maxdata);
    std::sort(f.pix.begin(), f.pix.end());
  }
  if (!std::isfinite(f.mindata) || !std::isfinite(f.maxdata))
    return fail(c, POFD_E_ARG,
                "calcLikeSingle::readDataFromFile: Invalid data range "
                "encountered in data file");
  s.fields.push_back(std::move(f));
  s.prepared = false;
  if (fid) *fid = static_cast<int>(s.fields.size() - 1);
  return POFD_OK;
}

extern "C" int pofd1d_prepare(pofd_ctx* c) {
  if (!c) return POFD_E_ARG;
  if (!c->configured)
    return fail(c, POFD_E_ARG, "pofd1d_prepare: pofd1d_configure not called");
  if (c->sets.empty()) return fail(c, POFD_E_ARG, "pofd1d_prepare: no beam sets");
  if (hipSetDevice(c->device) != hipSuccess)
    return fail(c, POFD_E_INTERNAL, "pofd1d_prepare: hipSetDevice failed");
  for (auto& s : c->sets) {
    s->minRFlux = s->maxRFlux = NAN;
    int rc = prepare_beamset(c, *s);
    if (rc) return rc;
  }
  return POFD_OK;
}

extern "C" int pofd1d_set_rrange(pofd_ctx* c, int b, double mn, double mx) {
  if (!c || b < 0 || static_cast<size_t>(b) >= c->sets.size())
    return fail(c, POFD_E_ARG, "pofd1d_set_rrange: bad beam set");
  Beamset& s = *c->sets[b];
  s.minRFlux = mn;
  s.maxRFlux = mx;
  return prepare_beamset(c, s);
}

extern "C" int pofd1d_beamset_info(const pofd_ctx* c, int b, double* out) {
  if (!c || !out || b < 0 || static_cast<size_t>(b) >= c->sets.size())
    return POFD_E_ARG;
  const Beamset& s = *c->sets[b];
  out[0] = s.minRFlux;
  out[1] = s.maxRFlux;
  out[2] = s.maxsigma_base;
  out[3] = s.minDataFlux;
  out[4] = s.maxDataFlux;
  out[5] = static_cast<double>(s.fields.size());
  out[6] = c->exp_conf;
  out[7] = s.prepared ? s.dev.dflux : NAN;
  return POFD_OK;
}

extern "C" int pofd1d_set_parts(pofd_ctx* c, int parts) {
  if (!c) return POFD_E_ARG;
  if (parts < 0 || parts > 4 || parts == 3)
    return fail(c, POFD_E_ARG, "pofd1d_set_parts: parts must be 0, 1, 2 or 4");
  c->force_parts = parts;
  ++c->generation;
  return POFD_OK;
}

namespace {
// d_count: nullptr, or the batch size on the device (at most nwalk);
// add_to: add to d_ll (and keep d_st's failures) instead of overwriting
int batch_device(pofd_ctx* c, const float* d_theta, uint32_t nwalk, const int32_t* d_count,
                 uint32_t stride, double* d_ll, int32_t* d_st, void* stream,
                 bool add_to = false) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (nwalk == 0) return POFD_OK;
  if (!d_theta || !d_ll || !d_st)
    return fail(c, POFD_E_ARG, "pofd1d_loglike_batch_device: null buffer");
  if (stride < c->knots.size() + 1)
    return fail(c, POFD_E_ARG,
                "calcLike::getLogLike: Not enough elements in paramSet");
  for (size_t b = 0; b < c->sets.size(); ++b) {
    Beamset& s = *c->sets[b];
    if ((rc = ensure_workspace(c, s, nwalk))) return rc;
    hipEvent_t* ev = nullptr;
    if ((rc = pofd2::ctx_stage_events(c, &ev))) return rc;
    s.dev.bonus = (b == 0) ? 1 : 0;  // the counts integrals once per batch
    rc = pofd::launch_loglike1d(
        s.dev, s.wk, static_cast<int>(b), d_theta, nwalk, d_count, stride, d_ll, d_st,
        (b > 0 || add_to) ? 1 : 0, c->force_parts, c->parts_tuned ? c->parts_tab.data() : nullptr,
        c->parts_tuned ? static_cast<const uint8_t*>(c->dparts->p) : nullptr, stream, ev);
    if (rc) return fail(c, rc, "pofd1d_loglike_batch_device: launch failed: " +
                                   std::string(hipGetErrorString(hipGetLastError())));
  }
  if (c->priors.flags || c->priors.reg_alpha > 0.0) {
    Beamset& s0 = *c->sets[0];
    rc = pofd::launch_prior1d(s0.dev, s0.wk, c->priors, d_theta, stride, nwalk, d_count, d_ll,
                              d_st, stream);
    if (rc) return fail(c, rc, "pofd1d_loglike_batch_device: prior launch failed");
  }
  c->last_nwalk = nwalk;
  return POFD_OK;
}
}  // namespace

// the counts integrals of the last batch (beam set 0's workspace; 4 per
// batch row), for the sampler's bonus parameters; null before any batch
const double* pofd1d_bonus_device(pofd_ctx* c) {
  if (!c || c->sets.empty() || !c->sets[0]->wbuf) return nullptr;
  return c->sets[0]->wk.bonus;
}

extern "C" int pofd1d_set_priors(pofd_ctx* c, int flags, double sigma_width, double cfirb_mean,
                                 double cfirb_sigma, double poisson_mean, double poisson_sigma,
                                 double reg_alpha) {
  if (!c) return POFD_E_ARG;
  if (flags & ~(pofd::kPriorSigma | pofd::kPriorCfirb | pofd::kPriorPoisson))
    return fail(c, POFD_E_ARG, "pofd1d_set_priors: unknown prior flag");
  if (((flags & pofd::kPriorSigma) && !(sigma_width > 0.0)) ||
      ((flags & pofd::kPriorCfirb) && !(cfirb_sigma > 0.0)) ||
      ((flags & pofd::kPriorPoisson) && !(poisson_sigma > 0.0)) || reg_alpha < 0.0)
    return fail(c, POFD_E_ARG, "pofd1d_set_priors: invalid prior width or alpha");
  c->priors = pofd::PriorDev{flags, sigma_width, cfirb_mean, cfirb_sigma, poisson_mean,
                             poisson_sigma, reg_alpha};
  ++c->generation;
  return POFD_OK;
}

extern "C" int pofd1d_loglike_batch_device(pofd_ctx* c, const float* d_theta,
                                           uint32_t nwalk, uint32_t stride,
                                           double* d_ll, int32_t* d_st,
                                           void* stream) {
  return batch_device(c, d_theta, nwalk, nullptr, stride, d_ll, d_st, stream);
}

extern "C" int pofd1d_loglike_batch_device_count(pofd_ctx* c, const float* d_theta,
                                                 uint32_t max_walkers, const int32_t* d_count,
                                                 uint32_t stride, double* d_ll, int32_t* d_st,
                                                 void* stream) {
  if (!d_count) return fail(c, POFD_E_ARG, "pofd1d_loglike_batch_device_count: null count");
  return batch_device(c, d_theta, max_walkers, d_count, stride, d_ll, d_st, stream);
}

// ---- the allocation-free batch: the same launches as
// pofd1d_loglike_batch_device, with every beam set's workspace laid out in
// the device scratch of noalloc.hip instead of memory the context
// allocates.  Every kernel writes each workspace value before it reads it,
// so nothing in the scratch has to start at zero.
extern "C" int pofd1d_loglike_batch_device_noalloc(pofd_ctx* c, const float* d_theta,
                                                   uint32_t nwalk, uint32_t stride,
                                                   double* d_ll, int32_t* d_st, void* stream) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (nwalk == 0) return POFD_OK;
  if (!d_theta || !d_ll || !d_st)
    return fail(c, POFD_E_ARG, "pofd1d_loglike_batch_device_noalloc: null buffer");
  if (stride < c->knots.size() + 1)
    return fail(c, POFD_E_ARG, "calcLike::getLogLike: Not enough elements in paramSet");
  if (c->noalloc_rc != POFD_OK)
    return fail(c, POFD_E_INTERNAL,
                "pofd1d_loglike_batch_device_noalloc: the device scratch could not be resolved");
  char* scratch = static_cast<char*>(pofd::noalloc_scratch(c->device));
  if (!scratch)
    return fail(c, POFD_E_INTERNAL, "pofd1d_loglike_batch_device_noalloc: no device scratch");
  const size_t nb = c->sets.size();
  std::vector<WorkDev> wks(nb);
  std::vector<size_t> offs(nb);
  size_t total = 0;
  for (size_t b = 0; b < nb; ++b) {
    offs[b] = total;
    const size_t bytes = workspace_layout(c->sets[b]->dev, nwalk, nullptr, wks[b]) * sizeof(double);
    total += (bytes + 255) & ~size_t(255);
  }
  if (total > pofd::kNoallocScratchBytes)
    return fail(c, POFD_E_ARG,
                "pofd1d_loglike_batch_device_noalloc: " + std::to_string(nwalk) +
                    " walkers need " + std::to_string(total >> 20) +
                    " MiB of workspace, above the " +
                    std::to_string(pofd::kNoallocScratchBytes >> 20) + " MiB device scratch");
  for (size_t b = 0; b < nb; ++b) {
    Beamset& s = *c->sets[b];
    workspace_layout(s.dev, nwalk, reinterpret_cast<double*>(scratch + offs[b]), wks[b]);
    s.dev.bonus = (b == 0) ? 1 : 0;  // the counts integrals once per batch
    rc = pofd::launch_loglike1d(
        s.dev, wks[b], static_cast<int>(b), d_theta, nwalk, nullptr, stride, d_ll, d_st,
        b > 0 ? 1 : 0, c->force_parts, c->parts_tuned ? c->parts_tab.data() : nullptr,
        c->parts_tuned ? static_cast<const uint8_t*>(c->dparts->p) : nullptr, stream, nullptr);
    if (rc) return fail(c, rc, "pofd1d_loglike_batch_device_noalloc: launch failed");
  }
  if (c->priors.flags || c->priors.reg_alpha > 0.0) {
    rc = pofd::launch_prior1d(c->sets[0]->dev, wks[0], c->priors, d_theta, stride, nwalk,
                              nullptr, d_ll, d_st, stream);
    if (rc) return fail(c, rc, "pofd1d_loglike_batch_device_noalloc: prior launch failed");
  }
  return POFD_OK;
}

// ---- autotuning of the workgroups per walker: every allowed split of each
// given batch size is timed with HIP events over `reps` batches after one
// warm-up batch, the fastest is kept, and sizes not measured take the
// choice of the nearest measured size (the smaller one on a tie)
extern "C" int pofd1d_autotune_parts(pofd_ctx* c, const float* theta, uint32_t stride,
                                     const int32_t* sizes, uint32_t nsizes, uint32_t reps,
                                     int32_t* chosen) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (nsizes == 0) {  // back to the size rule
    c->parts_tuned = false;
    c->parts_tab.fill(0);
    ++c->generation;
    return POFD_OK;
  }
  if (!theta || !sizes || !chosen || reps == 0)
    return fail(c, POFD_E_ARG, "pofd1d_autotune_parts: null buffer or zero reps");
  if (stride < c->knots.size() + 1)
    return fail(c, POFD_E_ARG, "pofd1d_autotune_parts: theta stride too small");
  if (c->stage_timing)
    return fail(c, POFD_E_ARG, "pofd1d_autotune_parts: stage timing must be off");
  uint32_t nmax = 0;
  for (uint32_t i = 0; i < nsizes; ++i) {
    if (sizes[i] < 1 || sizes[i] > static_cast<int32_t>(pofd::kMaxSlots))
      return fail(c, POFD_E_ARG, "pofd1d_autotune_parts: sizes must be in [1, 256]");
    nmax = std::max<uint32_t>(nmax, static_cast<uint32_t>(sizes[i]));
  }
  // workspaces of at least 64 walkers hold the full 256 (walker, part) slots
  for (auto& s : c->sets)
    if ((rc = ensure_workspace(c, *s, std::max<uint32_t>(nmax, 64)))) return rc;
  std::unique_ptr<DevBuf> dth, dll, dst;
  const size_t tb = static_cast<size_t>(nmax) * stride * sizeof(float);
  if ((rc = dev_alloc(c, dth, tb)) || (rc = dev_alloc(c, dll, nmax * sizeof(double))) ||
      (rc = dev_alloc(c, dst, nmax * sizeof(int32_t))))
    return rc;
  hipError_t e = hipMemcpy(dth->p, theta, tb, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(theta)");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return hip_fail(c, e, "hipEventCreate");
  }
  const int saved_force = c->force_parts;
  const bool saved_tuned = c->parts_tuned;
  c->parts_tuned = false;
  std::array<uint8_t, pofd::kMaxSlots + 1> tab{};
  const float* d_theta = static_cast<const float*>(dth->p);
  double* d_ll = static_cast<double*>(dll->p);
  int32_t* d_st = static_cast<int32_t*>(dst->p);
  for (uint32_t i = 0; i < nsizes && rc == POFD_OK; ++i) {
    const uint32_t n = static_cast<uint32_t>(sizes[i]);
    const int pmax = std::min<int>(4, static_cast<int>(pofd::kMaxSlots / n));
    int best = 1;
    float best_ms = 0.0f;
    for (int p = 1; p <= pmax && rc == POFD_OK; p <<= 1) {
      c->force_parts = p;
      if ((rc = batch_device(c, d_theta, n, nullptr, stride, d_ll, d_st, nullptr))) break;
      (void)hipEventRecord(e0, nullptr);
      for (uint32_t r = 0; r < reps && rc == POFD_OK; ++r)
        rc = batch_device(c, d_theta, n, nullptr, stride, d_ll, d_st, nullptr);
      (void)hipEventRecord(e1, nullptr);
      if ((e = hipEventSynchronize(e1)) != hipSuccess) {
        rc = hip_fail(c, e, "hipEventSynchronize");
        break;
      }
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (p == 1 || ms < best_ms) {
        best = p;
        best_ms = ms;
      }
    }
    tab[n] = static_cast<uint8_t>(best);
    chosen[i] = best;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->force_parts = saved_force;
  if (rc) {
    c->parts_tuned = saved_tuned;
    return rc;
  }
  // unmeasured sizes: the nearest measured one
  for (int n = 1; n <= static_cast<int>(pofd::kMaxSlots); ++n) {
    if (tab[n]) continue;
    int bestd = 1 << 30;
    uint8_t v = 0;
    for (uint32_t i = 0; i < nsizes; ++i) {
      const int d = std::abs(sizes[i] - n);
      if (d < bestd || (d == bestd && sizes[i] < n)) {
        bestd = d;
        v = tab[sizes[i]];
      }
    }
    tab[n] = v;
  }
  if ((rc = dev_alloc(c, c->dparts, tab.size()))) return rc;
  e = hipMemcpy(c->dparts->p, tab.data(), tab.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(parts table)");
  c->parts_tab = tab;
  c->parts_tuned = true;
  ++c->generation;
  return POFD_OK;
}

// ---- several contexts (fields of different nflux sharing theta) as one
// captured HIP graph: the contexts' batches in order on one stream, each
// adding its log likelihood to the previous ones'
struct pofd_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // the contexts captured and their generations at the capture
  std::vector<std::pair<pofd_ctx*, uint64_t>> ctxs;
  ~pofd_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

extern "C" int pofd1d_graph_create(pofd_ctx** ctxs, uint32_t nctx, const float* d_theta,
                                   uint32_t nwalk, uint32_t stride, double* d_ll,
                                   int32_t* d_st, void* stream, pofd_graph** out) {
  if (!out || !ctxs || nctx == 0 || nwalk == 0 || !stream) return POFD_E_ARG;
  *out = nullptr;
  for (uint32_t i = 0; i < nctx; ++i) {
    pofd_ctx* c = ctxs[i];
    int rc = check_ready(c);
    if (rc) return rc;
    if (c->stage_timing)
      return fail(c, POFD_E_ARG, "pofd1d_graph_create: stage timing must be off");
    // every workspace exists before the capture: nothing allocates inside it
    for (auto& s : c->sets)
      if ((rc = ensure_workspace(c, *s, nwalk))) return rc;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  auto* g = new pofd_graph();
  hipError_t he = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  int rc = (he == hipSuccess) ? POFD_OK : POFD_E_INTERNAL;
  for (uint32_t i = 0; rc == POFD_OK && i < nctx; ++i)
    rc = batch_device(ctxs[i], d_theta, nwalk, nullptr, stride, d_ll, d_st, stream, i > 0);
  if (he == hipSuccess) {
    const hipError_t he2 = hipStreamEndCapture(st, &g->graph);
    if (he2 != hipSuccess && rc == POFD_OK) rc = POFD_E_INTERNAL;
  }
  if (rc == POFD_OK &&
      hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
    rc = POFD_E_INTERNAL;
  if (rc) {
    delete g;
    return rc;
  }
  // recorded after the capture: the captured launches allocate nothing, so
  // the generations are those the graph's pointers belong to
  for (uint32_t i = 0; i < nctx; ++i) g->ctxs.emplace_back(ctxs[i], ctxs[i]->generation);
  *out = g;
  return POFD_OK;
}

extern "C" int pofd1d_graph_launch(pofd_graph* g, void* stream) {
  if (!g || !g->exec) return POFD_E_ARG;
  for (auto& cg : g->ctxs)
    if (cg.first->generation != cg.second)
      return fail(cg.first, POFD_E_ARG,
                  "pofd1d_graph_launch: a context's workspace, tables or launch "
                  "parameters changed after the capture; create the graph again");
  return hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)) == hipSuccess
             ? POFD_OK
             : POFD_E_INTERNAL;
}

extern "C" void pofd1d_graph_destroy(pofd_graph* g) { delete g; }


extern "C" int pofd1d_loglike_batch(pofd_ctx* c, const float* theta,
                                    uint32_t nwalk, uint32_t stride,
                                    double* ll, int32_t* st) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (nwalk == 0) return POFD_OK;
  if (!theta || !ll || !st)
    return fail(c, POFD_E_ARG, "pofd1d_loglike_batch: null buffer");
  const size_t tb = static_cast<size_t>(nwalk) * stride * sizeof(float);
  if ((rc = dev_alloc(c, c->dtheta, tb))) return rc;
  if ((rc = dev_alloc(c, c->dll, nwalk * sizeof(double)))) return rc;
  if ((rc = dev_alloc(c, c->dst, nwalk * sizeof(int32_t)))) return rc;
  hipError_t e = hipMemcpy(c->dtheta->p, theta, tb, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(theta)");
  rc = pofd1d_loglike_batch_device(c, static_cast<const float*>(c->dtheta->p),
                                   nwalk, stride,
                                   static_cast<double*>(c->dll->p),
                                   static_cast<int32_t*>(c->dst->p), nullptr);
  if (rc) return rc;
  e = hipMemcpy(ll, c->dll->p, nwalk * sizeof(double), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(loglike)");
  e = hipMemcpy(st, c->dst->p, nwalk * sizeof(int32_t), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(status)");
  return POFD_OK;
}

namespace {
int dbg_record(pofd_ctx* c, uint32_t walker, int b, int f, double* rec) {
  if (!c || b < 0 || static_cast<size_t>(b) >= c->sets.size())
    return fail(c, POFD_E_ARG, "debug accessor: bad beam set");
  Beamset& s = *c->sets[b];
  if (!s.wbuf || walker >= s.wcap || walker >= c->last_nwalk)
    return fail(c, POFD_E_ARG, "debug accessor: no such walker in last batch");
  if (f < 0 || f >= s.dev.nfields)
    return fail(c, POFD_E_ARG, "debug accessor: bad field");
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(c, e, "hipDeviceSynchronize");
  const double* src = s.wk.dbg + (static_cast<size_t>(walker) * s.dev.nfields + f) * pofd::kDbg;
  e = hipMemcpy(rec, src, pofd::kDbg * sizeof(double), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(debug)");
  return POFD_OK;
}
}  // namespace

extern "C" int pofd1d_get_debug(pofd_ctx* c, uint32_t walker, int b, int f,
                                double* info) {
  if (!info) return POFD_E_ARG;
  return dbg_record(c, walker, b, f, info);
}

extern "C" int pofd1d_get_pd(pofd_ctx* c, uint32_t walker, int b, int f,
                             int is_log, double* pd_out, double* minflux,
                             double* dflux) {
  double rec[pofd::kDbg];
  int rc = dbg_record(c, walker, b, f, rec);
  if (rc) return rc;
  if (!pd_out) return fail(c, POFD_E_ARG, "pofd1d_get_pd: null output");
  if (rec[15] != POFD_OK)
    return fail(c, static_cast<int>(rec[15]),
                "pofd1d_get_pd: the P(D) of this walker failed");
  Beamset& s = *c->sets[b];
  const int N = s.dev.N;
  std::unique_ptr<DevBuf> out;
  if ((rc = dev_alloc(c, out, N * sizeof(double)))) return rc;
  const double* pofd = s.wk.pofd + (static_cast<size_t>(walker) * s.dev.nfields + f) * s.wk.pofd_stride;
  const int already_log = rec[18] != 0.0;
  if (already_log && is_log <= 0)
    return fail(c, POFD_E_ARG,
                "pofd1d_get_pd: unbinned fields keep only the log2 P(D)");
  // is_log < 0: the clamped, unrotated, unnormalised transform output
  const int split = (is_log < 0) ? 0 : static_cast<int>(rec[0]);
  const double itot = (is_log < 0) ? 1.0 : rec[16];
  rc = pofd::launch_materialize_pd(pofd, N, split, itot, is_log > 0 ? 1 : 0,
                                   already_log, static_cast<double*>(out->p),
                                   nullptr);
  if (rc) return fail(c, rc, "pofd1d_get_pd: launch failed");
  hipError_t e = hipMemcpy(pd_out, out->p, N * sizeof(double), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(pd)");
  if (minflux) *minflux = rec[17];
  if (dflux) *dflux = s.dev.dflux;
  return POFD_OK;
}

extern "C" int pofd1d_get_spectrum(pofd_ctx* c, uint32_t walker, int b,
                                   double* out, uint32_t capacity,
                                   uint32_t* n) {
  double rec[pofd::kDbg];
  int rc = dbg_record(c, walker, b, 0, rec);
  if (rc) return rc;
  Beamset& s = *c->sets[b];
  const uint32_t cnt = static_cast<uint32_t>(rec[19]) + 1;
  if (n) *n = cnt;
  if (!out) return POFD_OK;
  if (capacity < cnt) return fail(c, POFD_E_ARG, "pofd1d_get_spectrum: capacity");
  const double2* src = s.wk.prtrans + static_cast<size_t>(walker) * s.wk.spec_stride;
  hipError_t e = hipMemcpy(out, src, cnt * sizeof(double2), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(spectrum)");
  return POFD_OK;
}

namespace {
int get_rinterp_seg(pofd_ctx* c, uint32_t walker, int b, int seg,
                    double* flux_out, double* log2r_out, const char* who) {
  double rec[pofd::kDbg];
  int rc = dbg_record(c, walker, b, 0, rec);
  if (rc) return rc;
  Beamset& s = *c->sets[b];
  if (seg == 0 ? !s.dev.haspos : s.dev.nseg < 2)
    return fail(c, POFD_E_ARG, std::string(who) + ": the beam has no " +
                                   (seg == 0 ? "positive" : "negative") + " pixels");
  const size_t n = static_cast<size_t>(s.dev.ninterp);
  const size_t off = static_cast<size_t>(seg) * n;
  hipError_t e;
  if (flux_out) {
    e = hipMemcpy(flux_out, s.dev.xnode + off, n * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(xnode)");
  }
  if (log2r_out) {
    e = hipMemcpy(log2r_out, s.wk.ynode + walker * s.wk.ynode_stride + off,
                  n * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(ynode)");
  }
  return POFD_OK;
}
}  // namespace

extern "C" int pofd1d_get_rinterp(pofd_ctx* c, uint32_t walker, int b,
                                  double* flux_out, double* log2r_out) {
  return get_rinterp_seg(c, walker, b, 0, flux_out, log2r_out, "pofd1d_get_rinterp");
}

extern "C" int pofd1d_get_rinterp_neg(pofd_ctx* c, uint32_t walker, int b,
                                      double* flux_out, double* log2r_out) {
  return get_rinterp_seg(c, walker, b, 1, flux_out, log2r_out, "pofd1d_get_rinterp_neg");
}

extern "C" int pofd1d_get_bonus(pofd_ctx* c, uint32_t nwalk, double* out) {
  if (!c || !out) return POFD_E_ARG;
  if (c->sets.empty() || !c->sets[0]->wbuf || nwalk > c->last_nwalk)
    return fail(c, POFD_E_ARG, "pofd1d_get_bonus: no such walkers in the last batch");
  const Beamset& s = *c->sets[0];
  std::vector<double> tmp(static_cast<size_t>(nwalk) * 4);
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess)
    e = hipMemcpy(tmp.data(), s.wk.bonus, tmp.size() * sizeof(double), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(bonus)");
  for (uint32_t w = 0; w < nwalk; ++w)
    for (int k = 0; k < 3; ++k) out[3 * w + k] = tmp[4 * static_cast<size_t>(w) + k];
  return POFD_OK;
}

extern "C" int pofd1d_stage_times(pofd_ctx* c, double* ms_out,
                                  uint32_t* nlaunch) {
  if (!c || !ms_out) return POFD_E_ARG;
  for (int i = 0; i < POFD_NSTAGES; ++i) ms_out[i] = 0.0;
  for (size_t l = 0; l < c->stage_used; ++l) {
    auto& ev = c->stage_events[l];
    hipError_t e = hipEventSynchronize(ev[POFD_NSTAGES]);
    if (e != hipSuccess) return hip_fail(c, e, "hipEventSynchronize");
    for (int i = 0; i < POFD_NSTAGES; ++i) {
      float ms = 0.0f;
      e = hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
      if (e != hipSuccess) return hip_fail(c, e, "hipEventElapsedTime");
      ms_out[i] += ms;
    }
  }
  if (nlaunch) *nlaunch = static_cast<uint32_t>(c->stage_used);
  c->stage_used = 0;
  return POFD_OK;
}

// ---- PD::getLogLike / PDDouble::getLogLike over a caller's P(D)
namespace {
int upload_bytes(pofd_ctx* c, std::unique_ptr<DevBuf>& b, const void* src, size_t bytes) {
  int rc = dev_alloc(c, b, bytes);
  if (rc) return rc;
  hipError_t e = hipMemcpy(b->p, src, bytes, hipMemcpyHostToDevice);
  return e == hipSuccess ? POFD_OK : hip_fail(c, e, "hipMemcpy(pd loglike input)");
}

// the P(D) with one padding row of 1.0 (log2 1 = 0): an interpolation at
// exactly the top edge reads one cell past the array with weight 0, as the
// reference's does (src/PDDouble.cc:1066, :1206)
int upload_pd(pofd_ctx* c, const double* pd, size_t n, size_t pad) {
  int rc = dev_alloc(c, c->pdl_pd, (n + pad) * sizeof(double));
  if (rc) return rc;
  std::vector<double> tail(pad, 1.0);
  hipError_t e = hipMemcpy(c->pdl_pd->p, pd, n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess && pad)
    e = hipMemcpy(static_cast<double*>(c->pdl_pd->p) + n, tail.data(), pad * sizeof(double),
                  hipMemcpyHostToDevice);
  return e == hipSuccess ? POFD_OK : hip_fail(c, e, "hipMemcpy(pd)");
}

int finish_pd_loglike(pofd_ctx* c, int rc, double* out) {
  if (rc) return fail(c, rc, "PD::getLogLike: launch failed");
  hipError_t e = hipMemcpy(out, static_cast<double*>(c->pdl_part->p) + pofd::pd_like_scratch_doubles() - 1,
                           sizeof(double), hipMemcpyDeviceToHost);
  return e == hipSuccess ? POFD_OK : hip_fail(c, e, "hipMemcpy(pd loglike)");
}
}  // namespace

extern "C" int pofd_pd_loglike(pofd_ctx* c, const double* pd, uint32_t n, double minflux,
                               double dflux, int is_log, const double* pix, uint64_t npix,
                               const uint32_t* bins, uint32_t nbins, double bincent0,
                               double bindelta, double* out) {
  if (!c || !out) return POFD_E_ARG;
  if (!pd || n < 2) return fail(c, POFD_E_ARG, "PD::getLogLike: pd not filled before likelihood calc");
  if ((bins && nbins == 0) || (!bins && (!pix || npix == 0)))
    return fail(c, POFD_E_ARG, "PD::getLogLike: No data present");
  if (!(dflux > 0.0)) return fail(c, POFD_E_ARG, "PD::getLogLike: dflux must be > 0");
  if (hipSetDevice(c->device) != hipSuccess)
    return fail(c, POFD_E_INTERNAL, "PD::getLogLike: hipSetDevice failed");
  int rc = upload_pd(c, pd, n, 2);
  if (!rc) rc = dev_alloc(c, c->pdl_part, pofd::pd_like_scratch_doubles() * sizeof(double));
  if (!rc && bins) rc = upload_bytes(c, c->pdl_bins, bins, nbins * sizeof(uint32_t));
  if (!rc && !bins) rc = upload_bytes(c, c->pdl_d1, pix, npix * sizeof(double));
  if (rc) return rc;
  rc = pofd::launch_pd_loglike1d(
      static_cast<const double*>(c->pdl_pd->p), static_cast<int>(n), minflux, dflux, is_log ? 1 : 0,
      bins ? nullptr : static_cast<const double*>(c->pdl_d1->p), bins ? 0 : npix,
      bins ? static_cast<const uint32_t*>(c->pdl_bins->p) : nullptr, bins ? nbins : 0, bincent0,
      bindelta, static_cast<double*>(c->pdl_part->p), nullptr);
  return finish_pd_loglike(c, rc, out);
}

extern "C" int pofd_pd2_loglike(pofd_ctx* c, const double* pd, uint32_t n1, uint32_t n2,
                                double minflux1, double dflux1, double minflux2, double dflux2,
                                int is_log, const double* pix1, const double* pix2,
                                uint64_t npix, const uint32_t* bins, uint32_t nbins1,
                                uint32_t nbins2, double bincent01, double bindelta1,
                                double bincent02, double bindelta2, double* out) {
  if (!c || !out) return POFD_E_ARG;
  if (!pd || n1 < 2 || n2 < 2)
    return fail(c, POFD_E_ARG, "PDDouble::getLogLike: pd not filled before likelihood calc");
  if ((bins && (nbins1 == 0 || nbins2 == 0)) || (!bins && (!pix1 || !pix2 || npix == 0)))
    return fail(c, POFD_E_ARG, "PDDouble::getLogLike: No data present");
  if (!(dflux1 > 0.0) || !(dflux2 > 0.0))
    return fail(c, POFD_E_ARG, "PDDouble::getLogLike: dflux must be > 0");
  if (hipSetDevice(c->device) != hipSuccess)
    return fail(c, POFD_E_INTERNAL, "PDDouble::getLogLike: hipSetDevice failed");
  const size_t nn = static_cast<size_t>(n1) * n2;
  int rc = upload_pd(c, pd, nn, static_cast<size_t>(n2) + 2);
  if (!rc) rc = dev_alloc(c, c->pdl_part, pofd::pd_like_scratch_doubles() * sizeof(double));
  if (!rc && bins)
    rc = upload_bytes(c, c->pdl_bins, bins, static_cast<size_t>(nbins1) * nbins2 * sizeof(uint32_t));
  if (!rc && !bins) rc = upload_bytes(c, c->pdl_d1, pix1, npix * sizeof(double));
  if (!rc && !bins) rc = upload_bytes(c, c->pdl_d2, pix2, npix * sizeof(double));
  if (rc) return rc;
  rc = pofd::launch_pd_loglike2d(
      static_cast<const double*>(c->pdl_pd->p), static_cast<int>(n1), static_cast<int>(n2), minflux1,
      dflux1, minflux2, dflux2, is_log ? 1 : 0,
      bins ? nullptr : static_cast<const double*>(c->pdl_d1->p),
      bins ? nullptr : static_cast<const double*>(c->pdl_d2->p), bins ? 0 : npix,
      bins ? static_cast<const uint32_t*>(c->pdl_bins->p) : nullptr, bins ? nbins1 : 0,
      bins ? nbins2 : 0, bincent01, bindelta1, bincent02, bindelta2,
      static_cast<double*>(c->pdl_part->p), nullptr);
  return finish_pd_loglike(c, rc, out);
}
