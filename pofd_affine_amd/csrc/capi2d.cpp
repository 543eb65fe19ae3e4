// C ABI of the two-band (2D) path: beams, data, context set-up and the
// batched launches.  The host part restates, once per fit, the
// theta-independent set-up of the reference's 2D likelihood:
//   doublebeam::setBeams / makeHistogram      src/doublebeam.cc:331-584
//   fitsDataDouble::applyBinning / meanSubtract
//                                             src/fitsDataDouble.cc:517-608
//   calcLikeDoubleSingle::setSigmaBase1/2     src/calcLikeDouble.cc:360-420
//   calcLikeDoubleSingle::setRRange           src/calcLikeDouble.cc:496-534
//   numberCountsDoubleLogNormal::getRRange with getMaxFlux and the
//     log-normal solver (GSL Brent)           src/numberCountsDoubleLogNormal.cc:
//                                             1352-1547
//   PDFactoryDouble::initRFlux and the edge grids of initR
//                                             src/PDFactoryDouble.cc:18-48,
//                                             313-345, 386-436
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "internal2d.h"

namespace {

using pofd2::kMaxK;

const double kPi2 = 3.141592653589793238462643383279502884197;

struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  Buf() = default;
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// doublebeam, host, bit-exact with the reference arithmetic
// ---------------------------------------------------------------------------
struct pofd_dbeam {
  double pixsize = 0.0, minval = 0.0;
  bool hassign[4] = {false, false, false, false};
  uint32_t npix[4] = {0, 0, 0, 0};
  std::vector<double> pix1[4], pix2[4], ipix1[4], ipix2[4], logratio[4];
  double tot1[4] = {}, tot2[4] = {}, totsq1[4] = {}, totsq2[4] = {};
  double min1[4] = {}, max1[4] = {}, min2[4] = {}, max2[4] = {};
  uint32_t nbins = 0;
  bool ishist[4] = {false, false, false, false};
  uint32_t nhist[4] = {0, 0, 0, 0};
  std::vector<double> binw[4], binv1[4], binv2[4], binlogratio[4];
};

namespace {

thread_local std::string g_dbeam_err;

// doublebeam::setBeams (src/doublebeam.cc:331-460) with getLogRatio
// (:722-748) formed eagerly
int dbeam_set(pofd_dbeam& bm, const double* beam1, const double* beam2, size_t n,
              double pixsize, double minval) {
  bm = pofd_dbeam();
  if (n == 0) {
    g_dbeam_err = "doublebeam::setBeams: Invalid (zero) n";
    return POFD_E_ARG;
  }
  if (minval < 0.0) {
    g_dbeam_err = "doublebeam::setBeams: Invalid (negative) MINVAL";
    return POFD_E_ARG;
  }
  bm.minval = minval;
  std::vector<unsigned> component(n);
  for (size_t i = 0; i < n; ++i) {
    const double v1 = beam1[i], v2 = beam2[i];
    if ((std::fabs(v1) > minval) && (std::fabs(v2) > minval)) {
      unsigned sgn;
      if (v1 > 0)
        sgn = (v2 > 0) ? 0 : 1;
      else
        sgn = (v2 > 0) ? 2 : 3;
      component[i] = sgn;
      bm.npix[sgn] += 1;
    } else {
      component[i] = 4;
    }
  }
  if (bm.npix[0] + bm.npix[1] + bm.npix[2] + bm.npix[3] == 0) {
    g_dbeam_err = "doublebeam::setBeams: No acceptable values found";
    return POFD_E_ARG;
  }
  for (int s = 0; s < 4; ++s) {
    const unsigned cn = bm.npix[s];
    if (cn == 0) {
      bm.min1[s] = bm.max1[s] = bm.min2[s] = bm.max2[s] = NAN;
      continue;
    }
    bm.hassign[s] = true;
    for (size_t j = 0; j < n; ++j)
      if (component[j] == static_cast<unsigned>(s)) {
        bm.pix1[s].push_back(std::fabs(beam1[j]));
        bm.pix2[s].push_back(std::fabs(beam2[j]));
      }
    const std::vector<double>& p1 = bm.pix1[s];
    const std::vector<double>& p2 = bm.pix2[s];
    double v = p1[0];
    bm.tot1[s] = v;
    bm.totsq1[s] = v * v;
    for (unsigned j = 1; j < cn; ++j) {
      v = p1[j];
      bm.tot1[s] += v;
      bm.totsq1[s] += v * v;
    }
    v = p2[0];
    bm.tot2[s] = v;
    bm.totsq2[s] = v * v;
    for (unsigned j = 1; j < cn; ++j) {
      v = p2[j];
      bm.tot2[s] += v;
      bm.totsq2[s] += v * v;
    }
    bm.ipix1[s].resize(cn);
    bm.ipix2[s].resize(cn);
    for (unsigned j = 0; j < cn; ++j) bm.ipix1[s][j] = 1.0 / p1[j];
    for (unsigned j = 0; j < cn; ++j) bm.ipix2[s][j] = 1.0 / p2[j];
    double mn = p1[0], mx = p1[0];
    for (unsigned j = 1; j < cn; ++j) {
      const double x = p1[j];
      if (x > mx)
        mx = x;
      else if (x < mn)
        mn = x;
    }
    bm.min1[s] = mn;
    bm.max1[s] = mx;
    mn = mx = p2[0];
    for (unsigned j = 1; j < cn; ++j) {
      const double x = p2[j];
      if (x > mx)
        mx = x;
      else if (x < mn)
        mn = x;
    }
    bm.min2[s] = mn;
    bm.max2[s] = mx;
    bm.logratio[s].resize(cn);
    for (unsigned j = 0; j < cn; ++j) bm.logratio[s][j] = std::log(p1[j] * bm.ipix2[s][j]);
  }
  if (pixsize <= 0.0) {
    g_dbeam_err = "doublebeam::setBeams: Invalid (non-positive) pixel size";
    return POFD_E_ARG;
  }
  bm.pixsize = pixsize;
  return POFD_OK;
}

// doublebeam::makeHistogram (src/doublebeam.cc:467-584), histothresh 30
int dbeam_hist(pofd_dbeam& bm, uint32_t NBINS) {
  if (NBINS == 0) {
    g_dbeam_err = "doublebeam::makeHistogram: Invalid (zero) NBINS";
    return POFD_E_ARG;
  }
  bm.nbins = NBINS;
  const size_t nb2 = static_cast<size_t>(NBINS) * NBINS;
  std::vector<unsigned> wt(nb2);
  std::vector<double> h1(nb2), h2(nb2);
  const double log2outscale = 0.0014419741739063218;
  const double dnbins = static_cast<double>(NBINS);
  for (int s = 0; s < 4; ++s) {
    bm.ishist[s] = false;
    bm.nhist[s] = 0;
    bm.binw[s].clear();
    bm.binv1[s].clear();
    bm.binv2[s].clear();
    bm.binlogratio[s].clear();
    if (!(bm.hassign[s] && bm.npix[s] >= 30)) continue;
    const double minb1 = std::log2(bm.min1[s]) - log2outscale;
    const double maxb1 = std::log2(bm.max1[s]) + log2outscale;
    const double ih1 = dnbins / (maxb1 - minb1);
    const double minb2 = std::log2(bm.min2[s]) - log2outscale;
    const double maxb2 = std::log2(bm.max2[s]) + log2outscale;
    const double ih2 = dnbins / (maxb2 - minb2);
    std::fill(wt.begin(), wt.end(), 0u);
    std::fill(h1.begin(), h1.end(), 0.0);
    std::fill(h2.begin(), h2.end(), 0.0);
    for (unsigned i = 0; i < bm.npix[s]; ++i) {
      const double v1 = bm.pix1[s][i], v2 = bm.pix2[s][i];
      const unsigned i1 = static_cast<unsigned>((std::log2(v1) - minb1) * ih1);
      const unsigned i2 = static_cast<unsigned>((std::log2(v2) - minb2) * ih2);
      const size_t t = static_cast<size_t>(i1) * NBINS + i2;
      wt[t] += 1;
      h1[t] += v1;
      h2[t] += v2;
    }
    for (size_t i = 0; i < nb2; ++i)
      if (wt[i] > 0) {
        const double wn = static_cast<double>(wt[i]);
        bm.binw[s].push_back(wn);
        bm.binv1[s].push_back(wn / h1[i]);
        bm.binv2[s].push_back(wn / h2[i]);
      }
    for (size_t i = 0; i < bm.binw[s].size(); ++i)
      bm.binlogratio[s].push_back(std::log(bm.binv2[s][i] / bm.binv1[s][i]));
    bm.ishist[s] = true;
    bm.nhist[s] = static_cast<uint32_t>(bm.binw[s].size());
  }
  return POFD_OK;
}

// ---------------------------------------------------------------------------
// numberCountsDoubleLogNormal, host side (R range only)
// ---------------------------------------------------------------------------
// GSL natural cubic spline (cspline_init + coeff_calc), coefficients
// y, b, c, d per interval
void cspline_host(int n, const double* x, const double* y, std::vector<double>& cf) {
  std::vector<double> c(n, 0.0), diag(n), off(n), g(n), alpha(n), gamma(n);
  const int sys = n - 2;
  for (int i = 0; i < sys; ++i) {
    const double h_i = x[i + 1] - x[i], h_ip1 = x[i + 2] - x[i + 1];
    const double yd_i = y[i + 1] - y[i], yd_ip1 = y[i + 2] - y[i + 1];
    const double g_i = (h_i != 0.0) ? 1.0 / h_i : 0.0;
    const double g_ip1 = (h_ip1 != 0.0) ? 1.0 / h_ip1 : 0.0;
    off[i] = h_ip1;
    diag[i] = 2.0 * (h_ip1 + h_i);
    g[i] = 3.0 * (yd_ip1 * g_ip1 - yd_i * g_i);
  }
  if (sys == 1) {
    c[1] = g[0] / diag[0];
  } else if (sys > 1) {
    std::vector<double>& z = diag;
    alpha[0] = diag[0];
    gamma[0] = off[0] / alpha[0];
    for (int i = 1; i < sys - 1; ++i) {
      alpha[i] = diag[i] - off[i - 1] * gamma[i - 1];
      gamma[i] = off[i] / alpha[i];
    }
    alpha[sys - 1] = diag[sys - 1] - off[sys - 2] * gamma[sys - 2];
    z[0] = g[0];
    for (int i = 1; i < sys; ++i) z[i] = g[i] - gamma[i - 1] * z[i - 1];
    for (int i = 0; i < sys; ++i) z[i] = z[i] / alpha[i];
    c[sys] = z[sys - 1];
    for (int i = sys - 2; i >= 0; --i) c[i + 1] = z[i] - gamma[i] * c[i + 2];
  }
  cf.assign(4 * (n - 1), 0.0);
  for (int i = 0; i < n - 1; ++i) {
    const double dx = x[i + 1] - x[i], dy = y[i + 1] - y[i];
    cf[4 * i] = y[i];
    cf[4 * i + 1] = dy / dx - dx * (c[i + 1] + 2.0 * c[i]) / 3.0;
    cf[4 * i + 2] = c[i];
    cf[4 * i + 3] = (c[i + 1] - c[i]) / (3.0 * dx);
  }
}

int bsearch_host(const double* x, int n, double v) {
  int lo = 0, hi = n - 1;
  while (hi > lo + 1) {
    const int m = (lo + hi) >> 1;
    if (x[m] > v)
      hi = m;
    else
      lo = m;
  }
  return lo;
}

struct LNHost {
  int nk = 0, ns = 0, no = 0;
  double knots[kMaxK] = {}, logknots[kMaxK] = {}, sk[kMaxK] = {}, ok[kMaxK] = {};
  double min_sigma = 0.0;
  std::vector<double> sv, ov, scf, ocf;
  bool valid = false;

  void set_params(const float* th) {
    valid = true;
    const double logfac = std::log2(10.0);
    for (int i = 0; i < nk; ++i)
      if (std::isnan(logfac * static_cast<double>(th[i]))) valid = false;
    sv.assign(ns, 0.0);
    ov.assign(no, 0.0);
    for (int i = 0; i < ns; ++i) {
      sv[i] = th[nk + i];
      if (std::isnan(sv[i]) || std::isinf(sv[i])) valid = false;
    }
    for (int i = 0; i < no; ++i) {
      ov[i] = th[nk + ns + i];
      if (std::isnan(ov[i])) valid = false;
    }
    if (min_sigma < 0.0) valid = false;
    if (valid && ns > 2) cspline_host(ns, sk, sv.data(), scf);
    if (valid && no > 2) cspline_host(no, ok, ov.data(), ocf);
  }
  double interp(const double* x, const std::vector<double>& vals, const std::vector<double>& cf,
                int n, double f1, bool floored) const {
    if (n == 1) return vals[0];
    if (f1 <= x[0]) return vals[0];
    if (f1 >= x[n - 1]) return vals[n - 1];
    double val;
    if (n == 2) {
      const double dx = x[1] - x[0];
      val = (dx > 0.0) ? vals[0] + (f1 - x[0]) / dx * (vals[1] - vals[0]) : 0.0;
    } else {
      const int k = bsearch_host(x, n, f1);
      const double dx = f1 - x[k];
      const double* c = cf.data() + 4 * k;
      val = c[0] + dx * (c[1] + dx * (c[2] + dx * c[3]));
    }
    if (floored && val < min_sigma) val = min_sigma;
    return val;
  }
  double sigma(double f1) const { return interp(sk, sv, scf, ns, f1, true); }
  double offset(double f1) const { return interp(ok, ov, ocf, no, f1, false); }
};

// the log-normal density minus the target (src/numberCountsDoubleLogNormal.cc:
// 2532-2546)
struct LNP {
  double mu, sig2, isig, isig2, target;
};
double lognorm(double u, const LNP& p) {
  const double prefac = 1.0 / std::sqrt(2 * kPi2);
  const double eu = std::exp(-u);
  const double arg = u - p.mu;
  const double L = prefac * std::exp(-0.5 * arg * arg * p.isig2) * p.isig * eu;
  return L - p.target;
}

// GSL roots/brent.c iteration and gsl_root_test_interval, as published
struct Brent {
  double a, b, c, d, e, fa, fb, fc;
};
void brent_iterate(Brent& s, const LNP& p, double& root, double& xlo, double& xhi) {
  double a = s.a, b = s.b, c = s.c, d = s.d, e = s.e, fa = s.fa, fb = s.fb, fc = s.fc;
  bool ac_equal = false;
  if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) {
    ac_equal = true;
    c = a;
    fc = fa;
    d = b - a;
    e = b - a;
  }
  if (std::fabs(fc) < std::fabs(fb)) {
    ac_equal = true;
    a = b;
    b = c;
    c = a;
    fa = fb;
    fb = fc;
    fc = fa;
  }
  const double tol = 0.5 * DBL_EPSILON * std::fabs(b);
  const double m = 0.5 * (c - b);
  if (fb == 0) {
    root = xlo = xhi = b;
    return;
  }
  if (std::fabs(m) <= tol) {
    root = b;
    xlo = std::min(b, c);
    xhi = std::max(b, c);
    return;
  }
  if (std::fabs(e) < tol || std::fabs(fa) <= std::fabs(fb)) {
    d = m;
    e = m;
  } else {
    double pp, q, r;
    const double ss = fb / fa;
    if (ac_equal) {
      pp = 2 * m * ss;
      q = 1 - ss;
    } else {
      q = fa / fc;
      r = fb / fc;
      pp = ss * (2 * m * q * (q - r) - (b - a) * (r - 1));
      q = (q - 1) * (r - 1) * (ss - 1);
    }
    if (pp > 0)
      q = -q;
    else
      pp = -pp;
    const double lim1 = 3 * m * q - std::fabs(tol * q), lim2 = std::fabs(e * q);
    if (2 * pp < std::min(lim1, lim2)) {
      e = d;
      d = pp / q;
    } else {
      d = m;
      e = m;
    }
  }
  a = b;
  fa = fb;
  if (std::fabs(d) > tol)
    b += d;
  else
    b += (m > 0 ? +tol : -tol);
  fb = lognorm(b, p);
  s = Brent{a, b, c, d, e, fa, fb, fc};
  root = b;
  if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) c = a;
  xlo = std::min(b, c);
  xhi = std::max(b, c);
}

// logNormalSolver (src/numberCountsDoubleLogNormal.cc:1389-1478)
int lognormal_solver(double mu, double sig, double nsig, double& out, std::string& err) {
  if (sig <= 0.0) {
    err = "numberCountsDoubleLogNormal::logNormalSolver: Invalid (non-positive) sigma";
    return POFD_E_INVALID_MODEL;
  }
  const double prefac = 1.0 / std::sqrt(2.0 * kPi2);
  const double sig2 = sig * sig, mode_u = mu - sig2;
  const double isig = 1.0 / sig;
  const double val_mode = prefac * isig * std::exp(0.5 * sig2 - mu);
  const LNP p{mu, sig2, isig, isig * isig, val_mode * std::exp(-0.5 * nsig * nsig)};
  double u_low = mode_u;
  double delta_u = std::log(std::exp(mode_u) + 2 * sig) - mode_u;
  double fval = lognorm(u_low + delta_u, p);
  if (fval >= 0.0) {
    for (int i = 0; i < 20; ++i) {
      delta_u *= 1.5;
      fval = lognorm(u_low + delta_u, p);
      if (fval < 0.0) break;
    }
    if (fval >= 0.0) {
      err = "numberCountsDoubleLogNormal::logNormalSolver: Unable to bracket Log Normal";
      return POFD_E_INVALID_MODEL;
    }
  }
  double u_high = u_low + delta_u;
  Brent s{u_low, u_high, u_high, u_high - u_low, u_high - u_low, lognorm(u_low, p),
          lognorm(u_high, p), 0.0};
  s.fc = s.fb;
  if ((s.fa < 0.0 && s.fb < 0.0) || (s.fa > 0.0 && s.fb > 0.0)) {
    err = "numberCountsDoubleLogNormal::logNormalSolver: endpoints do not straddle y=0";
    return POFD_E_INVALID_MODEL;
  }
  double u = 0.5 * (u_low + u_high);
  for (int i = 0; i < 20; ++i) {
    brent_iterate(s, p, u, u_low, u_high);
    const double alo = std::fabs(u_low), ahi = std::fabs(u_high);
    const double min_abs =
        ((u_low > 0.0 && u_high > 0.0) || (u_low < 0.0 && u_high < 0.0)) ? std::min(alo, ahi)
                                                                          : 0.0;
    if (std::fabs(u_high - u_low) < 1e-3 * min_abs) {
      out = std::exp(u);
      return POFD_OK;
    }
  }
  err = "numberCountsDoubleLogNormal::logNormalSolver: Unable to brent bracket Log Normal";
  return POFD_E_INVALID_MODEL;
}

// ---------------------------------------------------------------------------
// context state
// ---------------------------------------------------------------------------
struct Field2Host {
  std::vector<double> d1, d2;
  std::vector<uint32_t> bins;
  uint32_t nb1 = 0, nb2 = 0;
  double bininfo[4] = {};
  bool binned = false;
  uint64_t npix = 0;
  double sigma_base1 = 0, sigma_base2 = 0, like_norm = 1, like_offset = 0;
  double mn1 = 0, mx1 = 0, mn2 = 0, mx2 = 0;
  std::unique_ptr<Buf> dd1, dd2, dbins;
};

struct Set2 {
  pofd_dbeam bm;
  std::vector<Field2Host> fields;
  double range[4] = {NAN, NAN, NAN, NAN};
  double maxsig1 = NAN, maxsig2 = NAN;
  double mind1 = NAN, maxd1 = NAN, mind2 = NAN, maxd2 = NAN;
  bool prepared = false;
  pofd2::Set2Dev dev{};
  std::vector<std::unique_ptr<Buf>> tables;
  pofd2::RGrid gmain{}, g00{}, gx{}, gy{};
  pofd2::Work2 wk{};
  std::unique_ptr<Buf> wbuf;
  uint32_t wcap = 0;
};

}  // namespace

namespace pofd2 {
struct Ctx2 {
  bool configured = false;
  uint32_t n = 0, nedge = 0;
  int edge_integ = 1;
  LNHost model;
  double exp_conf1 = 0.0, exp_conf2 = 0.0;
  std::vector<std::unique_ptr<Set2>> sets;
  uint32_t last_nwalk = 0;
  std::unique_ptr<Buf> dtheta, dll, dst;
  Prior2Dev priors{};  // added on the device after the beam sets
};
void ctx2_free(Ctx2* c2) { delete c2; }
}  // namespace pofd2

namespace {

using pofd2::Ctx2;
using pofd2::ctx_fail;

int hipfail(pofd_ctx* c, hipError_t e, const char* what) {
  return ctx_fail(c, POFD_E_INTERNAL, std::string("pofd::hip: ") + what + ": " +
                                          hipGetErrorString(e));
}

int alloc(pofd_ctx* c, std::unique_ptr<Buf>& b, size_t bytes) {
  if (b && b->bytes >= bytes) return POFD_OK;
  b.reset(new Buf());
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(&b->p, bytes);
  if (e != hipSuccess) {
    b.reset();
    return hipfail(c, e, "hipMalloc");
  }
  b->bytes = bytes;
  return POFD_OK;
}

template <typename T>
int upload(pofd_ctx* c, Set2& s, const std::vector<T>& v, const T** out) {
  s.tables.emplace_back(new Buf());
  Buf& b = *s.tables.back();
  const size_t bytes = std::max<size_t>(v.size() * sizeof(T), 8);
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) return hipfail(c, e, "hipMalloc(table)");
  b.bytes = bytes;
  if (!v.empty()) {
    e = hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hipfail(c, e, "hipMemcpy(table)");
  }
  *out = static_cast<const T*>(b.p);
  return POFD_OK;
}

Ctx2* get2(pofd_ctx* c) {
  if (!c) return nullptr;
  Ctx2*& p = pofd2::ctx2_of(c);
  if (!p) p = new Ctx2();
  return p;
}

// initRFluxInternal (src/PDFactoryDouble.cc:18-48)
void init_rflux(unsigned n, double minflux, double maxflux, std::vector<double>& rflux,
                double& dflux) {
  rflux.assign(n, 0.0);
  const double inm1 = 1.0 / static_cast<double>(n - 1);
  dflux = (maxflux - minflux) * inm1;
  if (minflux >= 0.0) {
    rflux[0] = minflux;
    for (unsigned i = 1; i < n; ++i) rflux[i] = static_cast<double>(i) * dflux + minflux;
  } else {
    dflux = maxflux / (n - std::floor(-minflux / dflux) - 1.0);
    const unsigned w = static_cast<unsigned>(maxflux / dflux + 0.999999999);
    rflux[0] = 0.0;
    for (unsigned i = 1; i <= w; ++i) rflux[i] = static_cast<double>(i) * dflux;
    const double wrapval = -static_cast<double>(n) * dflux;
    for (unsigned i = w + 1; i < n; ++i) rflux[i] = static_cast<double>(i) * dflux + wrapval;
  }
}

std::vector<double2> tw_table(size_t L) {
  std::vector<double2> t(L);
  for (size_t k = 0; k < L; ++k) {
    const long double a = -2.0L * static_cast<long double>(kPi2) * static_cast<long double>(k) /
                          static_cast<long double>(L);
    t[k] = make_double2(static_cast<double>(cosl(a)), static_cast<double>(sinl(a)));
  }
  return t;
}

int prepare_set2(pofd_ctx* c, Ctx2& C2, Set2& s) {
  const unsigned n = C2.n, ne = C2.nedge;
  for (int k = 0; k < 4; ++k)
    if (std::isnan(s.range[k]))
      return ctx_fail(c, POFD_E_ARG, "calcLikeDoubleSingle::getNegLogLike: R range not set");
  double min1 = s.range[0], max1 = s.range[1], min2 = s.range[2], max2 = s.range[3];
  if (max1 <= 0.0 || max2 <= 0.0)
    return ctx_fail(c, POFD_E_ARG, "PDFactoryDouble::initPD: Invalid (non-positive) maxflux");
  if (max1 == min1 || max2 == min2)
    return ctx_fail(c, POFD_E_ARG, "PDFactoryDouble::initPD: Invalid range (0) between min/max flux");
  if (max1 < min1) std::swap(min1, max1);
  if (max2 < min2) std::swap(min2, max2);
  std::vector<double> rf1, rf2;
  double dflux1, dflux2;
  init_rflux(n, min1, max1, rf1, dflux1);
  init_rflux(n, min2, max2, rf2, dflux2);
  // edge grids (initR, :386-436): log-spaced in band 1, linear in band 2
  const double lowEdgeRMult = 1e-9;
  const double minedge1 = dflux1 * lowEdgeRMult, minedge2 = dflux2 * lowEdgeRMult;
  const double inedgem1 = 1.0 / static_cast<double>(ne - 1);
  const double dife1 = -std::log(lowEdgeRMult) * inedgem1;
  const double dife2 = (dflux2 - minedge2) * inedgem1;
  std::vector<double> edge1(ne), edge2(ne);
  for (unsigned i = 0; i < ne; ++i) edge1[i] = minedge1 * std::exp(static_cast<double>(i) * dife1);
  for (unsigned i = 0; i < ne; ++i) edge2[i] = minedge2 + static_cast<double>(i) * dife2;
  auto f2tables = [&](const std::vector<double>& f2, std::vector<double>& inv,
                      std::vector<double>& lg, std::vector<int>& sg) {
    const size_t m = f2.size();
    inv.assign(m, 0.0);
    lg.assign(m, 0.0);
    sg.assign(m, 0);
    for (size_t j = 0; j < m; ++j) {
      double v = f2[j];
      if (v == 0) continue;
      if (v < 0) {
        sg[j] = 1;
        v = std::fabs(v);
      }
      inv[j] = 1.0 / v;
      lg[j] = std::log(v);
    }
  };
  s.tables.clear();
  pofd2::Set2Dev& d = s.dev;
  std::memset(&d, 0, sizeof(d));
  d.n = static_cast<int>(n);
  d.nedge = static_cast<int>(ne);
  d.edge_integ = C2.edge_integ;
  d.nk = C2.model.nk;
  d.ns = C2.model.ns;
  d.no = C2.model.no;
  d.logfac = std::log2(10.0);
  for (int i = 0; i < kMaxK; ++i) {
    d.knots[i] = C2.model.knots[i];
    d.logknots[i] = C2.model.logknots[i];
    d.sigmaknots[i] = C2.model.sk[i];
    d.offsetknots[i] = C2.model.ok[i];
  }
  d.min_sigma = C2.model.min_sigma;
  d.maxsig1 = s.maxsig1;
  d.maxsig2 = s.maxsig2;
  const double normfac = 1.0 / std::sqrt(2 * kPi2);
  const double pixsizedeg = s.bm.pixsize / 3600.0;
  d.prefac = pixsizedeg * pixsizedeg * normfac;
  d.hasneg1 = (s.bm.hassign[2] || s.bm.hassign[3]) ? 1 : 0;
  d.dflux1 = dflux1;
  d.dflux2 = dflux2;
  d.iflux1 = (2.0 * kPi2) / (n * dflux1);
  d.iflux2 = (2.0 * kPi2) / (n * dflux2);
  d.iRxnorm = dife1 / (dflux1 - minedge1);
  d.iRynorm = dife2 / (dflux2 - minedge2);
  d.iR00norm = dife1 * dife2 / ((dflux1 - minedge1) * (dflux2 - minedge2));
  d.spectral_cut = pofd2::ctx_spectral_cut(c);
  // breakpoints of the counts integrals: log2 of the band-1 knots and of the
  // sigma and offset knots inside the band-1 range
  {
    const LNHost& lm = C2.model;
    std::vector<double> bp(lm.logknots, lm.logknots + lm.nk);
    const double lo = bp.front(), hi = bp.back();
    for (int i = 0; i < lm.ns; ++i) {
      const double v = log2(lm.sk[i]);
      if (v > lo && v < hi) bp.push_back(v);
    }
    for (int i = 0; i < lm.no; ++i) {
      const double v = log2(lm.ok[i]);
      if (v > lo && v < hi) bp.push_back(v);
    }
    std::sort(bp.begin(), bp.end());
    d.nbp = static_cast<int>(bp.size());
    for (size_t i = 0; i < bp.size(); ++i) d.bpts[i] = bp[i];
    pofd::gauss_legendre(32, d.glx, d.glw);
  }
  int rc;
  if ((rc = upload(c, s, rf1, &d.rflux1))) return rc;
  if ((rc = upload(c, s, rf2, &d.rflux2))) return rc;
  if ((rc = upload(c, s, edge1, &d.edge1))) return rc;
  std::vector<double2> th = tw_table(n / 2), tn = tw_table(n);
  if ((rc = upload(c, s, th, &d.tw_half))) return rc;
  if ((rc = upload(c, s, tn, &d.tw_n))) return rc;
  const double maxknot = C2.model.knots[C2.model.nk - 1];
  for (int k = 0; k < 4; ++k) {
    pofd2::BeamComp& bc = d.comp[k];
    bc.n = 0;
    if (!s.bm.hassign[k]) continue;
    if (s.bm.ishist[k]) {
      bc.n = static_cast<int>(s.bm.nhist[k]);
      if ((rc = upload(c, s, s.bm.binv1[k], &bc.ieta1))) return rc;
      if ((rc = upload(c, s, s.bm.binw[k], &bc.w))) return rc;
      if ((rc = upload(c, s, s.bm.binlogratio[k], &bc.logratio))) return rc;
    } else {
      bc.n = static_cast<int>(s.bm.npix[k]);
      if ((rc = upload(c, s, s.bm.ipix1[k], &bc.ieta1))) return rc;
      bc.w = nullptr;
      if ((rc = upload(c, s, s.bm.logratio[k], &bc.logratio))) return rc;
    }
    bc.maxf1 = maxknot * s.bm.max1[k];
  }
  // the R kernel stages 3 doubles per element of a sign pair in LDS
  for (int k = 0; k < 2; ++k) {
    const size_t pair = static_cast<size_t>(d.comp[2 * k].n) + d.comp[2 * k + 1].n;
    if (3 * pair * sizeof(double) > pofd2::kRgridMaxLds)
      return ctx_fail(c, POFD_E_ARG,
                      "pofd2d_prepare: beam has " + std::to_string(pair) +
                          " elements in one band-1 sign; histogram it (at most " +
                          std::to_string(pofd2::kRgridMaxLds / 24) + ")");
  }
  // flux-grid tables
  std::vector<double> inv_m, lg_m, inv_e, lg_e;
  std::vector<int> sg_m, sg_e;
  f2tables(rf2, inv_m, lg_m, sg_m);
  f2tables(edge2, inv_e, lg_e, sg_e);
  const double *d_inv_m, *d_lg_m, *d_inv_e, *d_lg_e, *d_e1;
  const int *d_sg_m, *d_sg_e;
  if ((rc = upload(c, s, inv_m, &d_inv_m))) return rc;
  if ((rc = upload(c, s, lg_m, &d_lg_m))) return rc;
  if ((rc = upload(c, s, sg_m, &d_sg_m))) return rc;
  if ((rc = upload(c, s, inv_e, &d_inv_e))) return rc;
  if ((rc = upload(c, s, lg_e, &d_lg_e))) return rc;
  if ((rc = upload(c, s, sg_e, &d_sg_e))) return rc;
  d_e1 = d.edge1;
  s.gmain = pofd2::RGrid{static_cast<int>(n), static_cast<int>(n), d.rflux1, d_inv_m, d_lg_m,
                         d_sg_m, nullptr, 0, static_cast<int>(n), dflux1 * dflux2};
  s.g00 = pofd2::RGrid{static_cast<int>(ne), static_cast<int>(ne), d_e1, d_inv_e, d_lg_e,
                       d_sg_e, nullptr, 0, static_cast<int>(ne), 1.0};
  s.gx = pofd2::RGrid{static_cast<int>(ne), static_cast<int>(n), d_e1, d_inv_m, d_lg_m,
                      d_sg_m, nullptr, 0, static_cast<int>(n), 1.0};
  s.gy = pofd2::RGrid{static_cast<int>(n), static_cast<int>(ne), d.rflux1, d_inv_e, d_lg_e,
                      d_sg_e, nullptr, 0, static_cast<int>(ne), 1.0};
  // datasets
  d.nfields = static_cast<int>(s.fields.size());
  for (size_t f = 0; f < s.fields.size(); ++f) {
    Field2Host& fh = s.fields[f];
    pofd2::Field2Dev& fd = d.fields[f];
    fd.sigma_base1 = fh.sigma_base1;
    fd.sigma_base2 = fh.sigma_base2;
    fd.like_offset = fh.like_offset;
    fd.like_norm = fh.like_norm;
    fd.binned = fh.binned ? 1 : 0;
    fd.nb1 = fh.nb1;
    fd.nb2 = fh.nb2;
    fd.bincent01 = fh.bininfo[0];
    fd.bindelta1 = fh.bininfo[1];
    fd.bincent02 = fh.bininfo[2];
    fd.bindelta2 = fh.bininfo[3];
    fd.npix = fh.binned ? 0 : fh.npix;
    if (fh.binned) {
      if ((rc = alloc(c, fh.dbins, fh.bins.size() * sizeof(uint32_t)))) return rc;
      hipError_t e = hipMemcpy(fh.dbins->p, fh.bins.data(), fh.bins.size() * sizeof(uint32_t),
                               hipMemcpyHostToDevice);
      if (e != hipSuccess) return hipfail(c, e, "hipMemcpy(bins)");
      fd.bins = static_cast<const uint32_t*>(fh.dbins->p);
    } else {
      if ((rc = alloc(c, fh.dd1, fh.npix * sizeof(double)))) return rc;
      if ((rc = alloc(c, fh.dd2, fh.npix * sizeof(double)))) return rc;
      hipError_t e = hipMemcpy(fh.dd1->p, fh.d1.data(), fh.npix * sizeof(double),
                               hipMemcpyHostToDevice);
      if (e == hipSuccess)
        e = hipMemcpy(fh.dd2->p, fh.d2.data(), fh.npix * sizeof(double), hipMemcpyHostToDevice);
      if (e != hipSuccess) return hipfail(c, e, "hipMemcpy(pixels)");
      fd.d1 = static_cast<const double*>(fh.dd1->p);
      fd.d2 = static_cast<const double*>(fh.dd2->p);
    }
  }
  s.wcap = 0;
  s.prepared = true;
  return POFD_OK;
}

int ensure_work2(pofd_ctx* c, Ctx2& C2, Set2& s, uint32_t nwalk) {
  pofd2::Work2& w = s.wk;
  w.nwalk = static_cast<int>(nwalk);
  if (s.wcap >= nwalk && s.wbuf) return POFD_OK;
  const size_t n = C2.n, ne = C2.nedge, nf = s.fields.size(), W = nwalk;
  const size_t nc = n / 2 + 1;
  auto r8 = [](size_t x) { return (x + 7) & ~size_t(7); };
  w.R_stride = r8(n * n);
  w.ex_stride = r8(ne * n);        // nedge x n (ex) or n x nedge (ey)
  w.edge_stride = r8(std::max(n, ne));
  w.ystride = static_cast<int>(nc);
  w.Y_stride = r8(n * nc);       // complex
  w.YT_stride = r8(nc * n);
  w.PR_stride = r8(nc * n);
  w.Q_stride = w.YT_stride;      // Q reuses YT (consumed by the forward)
  w.QT_stride = w.Y_stride;      // QT reuses Y
  w.P_stride = r8(n * n);
  const size_t nparts = 64;
  size_t tot = 0;
  const size_t o_model = tot; tot += r8(pofd2::kModelStride) * W;
  const size_t o_stats = tot; tot += r8(pofd2::kStats) * W;
  const size_t o_R = tot; tot += w.R_stride * W;
  const size_t o_e00 = tot; tot += r8(ne * ne) * W;
  const size_t o_ex = tot; tot += w.ex_stride * W;
  const size_t o_ey = tot; tot += w.ex_stride * W;
  const size_t o_rx = tot; tot += w.edge_stride * W;
  const size_t o_ry = tot; tot += w.edge_stride * W;
  const size_t o_r00 = tot; tot += w.edge_stride * W;
  const size_t o_Y = tot; tot += 2 * w.Y_stride * W;
  const size_t o_YT = tot; tot += 2 * w.YT_stride * W;
  const size_t o_PR = tot; tot += 2 * w.PR_stride * W;
  const size_t o_P = tot; tot += w.P_stride * nf * W;
  const size_t o_rs1 = tot; tot += w.edge_stride * nf * W;
  const size_t o_rs2 = tot; tot += w.edge_stride * nf * W;
  const size_t o_rA = tot; tot += w.edge_stride * nf * W;
  const size_t o_rB = tot; tot += w.edge_stride * nf * W;
  w.colpart_stride = r8(static_cast<size_t>(pofd2::kRowBlocks2) * n);
  const size_t o_cp = tot; tot += w.colpart_stride * W;
  const size_t o_rsp = tot; tot += r8(nparts * 4) * W;
  const size_t o_lp = tot; tot += r8(pofd2::kLikeBlocks) * nf * W;
  const size_t o_dbg = tot; tot += r8(pofd2::kDbg2) * nf * W;
  const size_t o_bo = tot; tot += 8 * W;
  // the R grids' per-row beam pieces (rows of the largest grid, n or nedge)
  int maxc = 1;
  for (int k = 0; k < 2; ++k)
    maxc = std::max(maxc, s.dev.comp[2 * k].n + s.dev.comp[2 * k + 1].n);
  w.rpieces_w = 3 * maxc;
  w.rpieces_stride = r8(std::max(n, ne) * static_cast<size_t>(w.rpieces_w));
  const size_t o_rp = tot; tot += w.rpieces_stride * W;
  s.wbuf.reset();
  int rc = alloc(c, s.wbuf, tot * sizeof(double));
  if (rc) return rc;
  double* b = static_cast<double*>(s.wbuf->p);
  w.model = b + o_model;
  w.stats = b + o_stats;
  w.R = b + o_R;
  w.e00 = b + o_e00;
  w.ex = b + o_ex;
  w.ey = b + o_ey;
  w.rx = b + o_rx;
  w.ry = b + o_ry;
  w.r00rows = b + o_r00;
  w.Y = reinterpret_cast<double2*>(b + o_Y);
  w.YT = reinterpret_cast<double2*>(b + o_YT);
  w.PR = reinterpret_cast<double2*>(b + o_PR);
  w.Q = w.YT;
  w.QT = w.Y;
  w.P = b + o_P;
  w.rsum1 = b + o_rs1;
  w.rsum2 = b + o_rs2;
  w.rowA = b + o_rA;
  w.rowB = b + o_rB;
  w.colpart = b + o_cp;
  w.rstat_part = b + o_rsp;
  w.lpart = b + o_lp;
  w.dbg = b + o_dbg;
  w.bonus = b + o_bo;
  w.rpieces = b + o_rp;
  s.gmain.out = w.R;
  s.gmain.wstride = w.R_stride;
  s.g00.out = w.e00;
  s.g00.wstride = r8(ne * ne);
  s.gx.out = w.ex;
  s.gx.wstride = w.ex_stride;
  s.gy.out = w.ey;
  s.gy.wstride = w.ex_stride;
  s.wcap = nwalk;
  // zero once so that debug reads of never-written slots are defined
  hipError_t e = hipMemset(s.wbuf->p, 0, tot * sizeof(double));
  if (e != hipSuccess) return hipfail(c, e, "hipMemset(workspace)");
  return POFD_OK;
}

int check_ready2(pofd_ctx* c, Ctx2* C2) {
  if (!C2 || !C2->configured)
    return ctx_fail(c, POFD_E_ARG, "calcLikeDouble::getLogLike: pofd2d_configure not called");
  if (C2->sets.empty()) return ctx_fail(c, POFD_E_ARG, "calcLikeDouble::getLogLike: no beam sets");
  for (auto& s : C2->sets)
    if (!s->prepared)
      return ctx_fail(c, POFD_E_ARG,
                      "calcLikeDoubleSingle::getNegLogLike: R range not set (pofd2d_prepare)");
  return POFD_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int pofd_dbeam_create(pofd_dbeam** out, const double* beam1, const double* beam2, uint64_t n,
                      double pixsize, double minval) {
  if (!out || !beam1 || !beam2) return POFD_E_ARG;
  std::unique_ptr<pofd_dbeam> b(new pofd_dbeam());
  int rc = dbeam_set(*b, beam1, beam2, n, pixsize, minval);
  if (rc) return rc;
  *out = b.release();
  return POFD_OK;
}

int pofd_dbeam_make_histogram(pofd_dbeam* b, uint32_t nbins) {
  if (!b) return POFD_E_ARG;
  return dbeam_hist(*b, nbins);
}

void pofd_dbeam_destroy(pofd_dbeam* b) { delete b; }

const char* pofd_dbeam_last_error(void) { return g_dbeam_err.c_str(); }

int pofd_dbeam_info(const pofd_dbeam* b, double* out) {
  if (!b || !out) return POFD_E_ARG;
  for (int s = 0; s < 4; ++s) {
    double* o = out + 12 * s;
    o[0] = b->hassign[s];
    o[1] = b->npix[s];
    o[2] = b->ishist[s];
    o[3] = b->nhist[s];
    o[4] = b->hassign[s] ? b->tot1[s] : 0.0;
    o[5] = b->hassign[s] ? b->tot2[s] : 0.0;
    o[6] = b->hassign[s] ? b->totsq1[s] : 0.0;
    o[7] = b->hassign[s] ? b->totsq2[s] : 0.0;
    o[8] = b->min1[s];
    o[9] = b->max1[s];
    o[10] = b->min2[s];
    o[11] = b->max2[s];
  }
  out[48] = b->pixsize;
  out[49] = b->minval;
  out[50] = b->nbins;
  return POFD_OK;
}

int pofd_dbeam_get(const pofd_dbeam* b, int sgn, int which, double* out, uint32_t* n) {
  if (!b || sgn < 0 || sgn > 3 || which < 0 || which > 8) return POFD_E_ARG;
  const std::vector<double>* v[9] = {&b->pix1[sgn],   &b->pix2[sgn],  &b->ipix1[sgn],
                                     &b->ipix2[sgn],  &b->logratio[sgn], &b->binw[sgn],
                                     &b->binv1[sgn],  &b->binv2[sgn], &b->binlogratio[sgn]};
  const std::vector<double>& a = *v[which];
  if (n) *n = static_cast<uint32_t>(a.size());
  if (out && !a.empty()) std::memcpy(out, a.data(), a.size() * sizeof(double));
  return POFD_OK;
}

int pofd_data2_bin(const double* d1, const double* d2, uint64_t n, uint32_t nb1, uint32_t nb2,
                   uint32_t* bins, double* info) {
  if (!d1 || !d2 || !bins || !info || n == 0 || nb1 == 0 || nb2 == 0) return POFD_E_ARG;
  double mn1 = d1[0], mx1 = d1[0], mn2 = d2[0], mx2 = d2[0];
  for (uint64_t i = 1; i < n; ++i) {
    if (d1[i] < mn1) mn1 = d1[i];
    if (d1[i] > mx1) mx1 = d1[i];
  }
  for (uint64_t i = 1; i < n; ++i) {
    if (d2[i] < mn2) mn2 = d2[i];
    if (d2[i] > mx2) mx2 = d2[i];
  }
  if (!std::isfinite(mn1) || !std::isfinite(mx1) || !std::isfinite(mn2) || !std::isfinite(mx2))
    return POFD_E_ARG;
  std::memset(bins, 0, sizeof(uint32_t) * nb1 * nb2);
  const double bd1 = (nb1 == 1) ? 2 * (mx1 - mn1) : (mx1 - mn1) / static_cast<double>(nb1 - 1);
  const double bd2 = (nb2 == 1) ? 2 * (mx2 - mn2) : (mx2 - mn2) / static_cast<double>(nb2 - 1);
  const double ib1 = 1.0 / bd1, ib2 = 1.0 / bd2;
  for (uint64_t i = 0; i < n; ++i) {
    const unsigned i1 = static_cast<unsigned>((d1[i] - mn1) * ib1 + 0.5);
    const unsigned i2 = static_cast<unsigned>((d2[i] - mn2) * ib2 + 0.5);
    if (i1 >= nb1 || i2 >= nb2) return POFD_E_ARG;
    bins[static_cast<size_t>(i1) * nb2 + i2] += 1;
  }
  info[0] = mn1;
  info[1] = bd1;
  info[2] = mn2;
  info[3] = bd2;
  return POFD_OK;
}

int pofd_data2_mean_subtract(double* d1, double* d2, uint64_t n, double* mean) {
  if (!d1 || !d2 || n == 0) return POFD_E_ARG;
  double m1 = d1[0], m2 = d2[0];
  for (uint64_t i = 1; i < n; ++i) m1 += d1[i];
  m1 /= static_cast<double>(n);
  for (uint64_t i = 1; i < n; ++i) m2 += d2[i];
  m2 /= static_cast<double>(n);
  for (uint64_t i = 0; i < n; ++i) d1[i] -= m1;
  for (uint64_t i = 0; i < n; ++i) d2[i] -= m2;
  if (mean) {
    mean[0] = m1;
    mean[1] = m2;
  }
  return POFD_OK;
}

int pofd2d_configure(pofd_ctx* c, uint32_t n, uint32_t nedge, int edge_integ,
                     const double* knots, uint32_t nknots, const double* sigmapos,
                     uint32_t nsigma, const double* offsetpos, uint32_t noffset,
                     double min_sigma, double exp_conf1, double exp_conf2) {
  Ctx2* C2 = get2(c);
  if (!C2) return POFD_E_ARG;
  if (n < 64 || n > 4096 || (n & (n - 1)))
    return ctx_fail(c, POFD_E_ARG, "pofd2d_configure: n must be a power of two in [64, 4096]");
  if (edge_integ && (nedge < 2 || nedge > n))
    return ctx_fail(c, POFD_E_ARG, "PDFactoryDouble::initR: Nedge is less than R fill size");
  if (!knots || nknots <= 2 || nknots > static_cast<uint32_t>(kMaxK))
    return ctx_fail(c, POFD_E_ARG, "numberCountsDoubleLogNormal: nknots must be > 2 (and <= 32)");
  if (!sigmapos || nsigma < 1 || nsigma > static_cast<uint32_t>(kMaxK) || !offsetpos ||
      noffset < 1 || noffset > static_cast<uint32_t>(kMaxK))
    return ctx_fail(c, POFD_E_ARG, "pofd2d_configure: 1 to 32 sigma and offset knots required");
  for (uint32_t i = 0; i < nknots; ++i)
    if (!(knots[i] > 0.0) || (i > 0 && knots[i] <= knots[i - 1]))
      return ctx_fail(c, POFD_E_ARG,
                      "numberCountsDoubleLogNormal::setKnotPositions: knots must be positive and increasing");
  for (uint32_t i = 0; i < nsigma; ++i)
    if (!(sigmapos[i] > 0.0) || (i > 0 && sigmapos[i] <= sigmapos[i - 1]))
      return ctx_fail(c, POFD_E_ARG, "numberCountsDoubleLogNormal::setSigmaPositions: invalid positions");
  for (uint32_t i = 0; i < noffset; ++i)
    if (!(offsetpos[i] > 0.0) || (i > 0 && offsetpos[i] <= offsetpos[i - 1]))
      return ctx_fail(c, POFD_E_ARG, "numberCountsDoubleLogNormal::setOffsetPositions: invalid positions");
  if (min_sigma < 0.0)
    return ctx_fail(c, POFD_E_ARG, "numberCountsDoubleLogNormal::setMinSigma: Invalid (negative) min sigma");
  C2->n = n;
  C2->nedge = nedge;
  C2->edge_integ = edge_integ ? 1 : 0;
  LNHost& m = C2->model;
  m = LNHost();
  m.nk = static_cast<int>(nknots);
  m.ns = static_cast<int>(nsigma);
  m.no = static_cast<int>(noffset);
  for (uint32_t i = 0; i < nknots; ++i) {
    m.knots[i] = knots[i];
    m.logknots[i] = std::log2(knots[i]);
  }
  for (uint32_t i = 0; i < nsigma; ++i) m.sk[i] = sigmapos[i];
  for (uint32_t i = 0; i < noffset; ++i) m.ok[i] = offsetpos[i];
  m.min_sigma = min_sigma;
  C2->exp_conf1 = exp_conf1;
  C2->exp_conf2 = exp_conf2;
  C2->configured = true;
  for (auto& s : C2->sets) s->prepared = false;
  return POFD_OK;
}

int pofd2d_add_beamset(pofd_ctx* c, const pofd_dbeam* bm, int* id) {
  Ctx2* C2 = get2(c);
  if (!C2 || !bm) return POFD_E_ARG;
  std::unique_ptr<Set2> s(new Set2());
  s->bm = *bm;
  C2->sets.push_back(std::move(s));
  if (id) *id = static_cast<int>(C2->sets.size() - 1);
  return POFD_OK;
}

int pofd2d_add_field(pofd_ctx* c, int b, const double* d1, const double* d2, uint64_t npix,
                     const uint32_t* bins, uint32_t nb1, uint32_t nb2, const double* bininfo,
                     double sigma_base1, double sigma_base2, double like_norm, int* fid) {
  Ctx2* C2 = get2(c);
  if (!C2 || b < 0 || static_cast<size_t>(b) >= C2->sets.size())
    return ctx_fail(c, POFD_E_ARG, "pofd2d_add_field: bad beam set");
  Set2& s = *C2->sets[b];
  if (s.fields.size() >= static_cast<size_t>(pofd2::kMaxFields2))
    return ctx_fail(c, POFD_E_ARG, "pofd2d_add_field: too many datasets");
  if (npix == 0 || !d1 || !d2)
    return ctx_fail(c, POFD_E_ARG, "fitsDataDouble::readData: No data");
  if (!(like_norm != 0.0)) return ctx_fail(c, POFD_E_ARG, "pofd2d_add_field: like_norm must be non-zero");
  Field2Host f;
  f.npix = npix;
  f.sigma_base1 = sigma_base1;
  f.sigma_base2 = sigma_base2;
  f.like_norm = like_norm;
  for (uint64_t i = 0; i < npix; ++i)
    if (!std::isfinite(d1[i]) || !std::isfinite(d2[i]))
      return ctx_fail(c, POFD_E_ARG, "pofd2d_add_field: non-finite pixel value");
  if (bins) {
    if (nb1 == 0 || nb2 == 0 || !bininfo)
      return ctx_fail(c, POFD_E_ARG, "fitsDataDouble::applyBinning: no bins");
    f.binned = true;
    f.nb1 = nb1;
    f.nb2 = nb2;
    f.bins.assign(bins, bins + static_cast<size_t>(nb1) * nb2);
    for (int k = 0; k < 4; ++k) f.bininfo[k] = bininfo[k];
    // fitsDataDouble::getMinMax binned branch (:488-494)
    f.mn1 = bininfo[0];
    f.mx1 = bininfo[0] + static_cast<double>(nb1 - 1) * bininfo[1];
    f.mn2 = bininfo[2];
    f.mx2 = bininfo[2] + static_cast<double>(nb2 - 1) * bininfo[3];
  } else {
    f.d1.assign(d1, d1 + npix);
    f.d2.assign(d2, d2 + npix);
    f.mn1 = f.mx1 = d1[0];
    f.mn2 = f.mx2 = d2[0];
    for (uint64_t i = 1; i < npix; ++i) {
      if (d1[i] < f.mn1) f.mn1 = d1[i];
      if (d1[i] > f.mx1) f.mx1 = d1[i];
      if (d2[i] < f.mn2) f.mn2 = d2[i];
      if (d2[i] > f.mx2) f.mx2 = d2[i];
    }
    // The log likelihood is a sum over pixel pairs, so their order is free
    // (up to rounding): store them in strips of band-1 flux, ascending in
    // band 2 inside a strip, so that neighbouring lanes interpolate in
    // neighbouring cells of the P(D).
    const double span1 = f.mx1 - f.mn1;
    const double sfac = (span1 > 0) ? 512.0 / span1 : 0.0;
    std::vector<uint32_t> strip(npix);
    std::vector<uint64_t> order(npix);
    for (uint64_t i = 0; i < npix; ++i) {
      const double sv = (d1[i] - f.mn1) * sfac;
      strip[i] = static_cast<uint32_t>(sv < 511.0 ? sv : 511.0);
      order[i] = i;
    }
    std::sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) {
      if (strip[a] != strip[b]) return strip[a] < strip[b];
      if (d2[a] != d2[b]) return d2[a] < d2[b];
      return a < b;
    });
    for (uint64_t i = 0; i < npix; ++i) {
      f.d1[i] = d1[order[i]];
      f.d2[i] = d2[order[i]];
    }
  }
  s.fields.push_back(std::move(f));
  s.prepared = false;
  if (fid) *fid = static_cast<int>(s.fields.size() - 1);
  return POFD_OK;
}

// calcLikeDouble::setRRanges(p) (src/calcLikeDouble.cc:1093-1097) with the
// sigma bases and like offsets of every dataset, then the device tables
int pofd2d_prepare(pofd_ctx* c, const float* theta, uint32_t npar) {
  Ctx2* C2 = get2(c);
  if (!C2 || !C2->configured)
    return ctx_fail(c, POFD_E_ARG, "pofd2d_prepare: pofd2d_configure not called");
  if (C2->sets.empty()) return ctx_fail(c, POFD_E_ARG, "pofd2d_prepare: no beam sets");
  LNHost& m = C2->model;
  if (!theta || npar < static_cast<uint32_t>(m.nk + m.ns + m.no))
    return ctx_fail(c, POFD_E_ARG, "numberCountsDoubleLogNormal::setParams: Not enough parameters present to set");
  if (hipSetDevice(pofd2::ctx_device(c)) != hipSuccess)
    return ctx_fail(c, POFD_E_INTERNAL, "pofd2d_prepare: hipSetDevice failed");
  m.set_params(theta);
  if (!m.valid)
    return ctx_fail(c, POFD_E_INVALID_MODEL, "numberCountsDoubleLogNormal::getRRange: Model is not valid");
  // getMaxFlux (:1365-1380)
  const double kv = m.knots[m.nk - 1];
  double m2 = NAN;
  std::string err;
  int rc = lognormal_solver(m.offset(kv), m.sigma(kv), 3.0, m2, err);
  if (rc) return ctx_fail(c, rc, err);
  const double mx1 = kv, mx2 = kv * m2;
  for (auto& sp : C2->sets) {
    Set2& s = *sp;
    if (s.fields.empty()) return ctx_fail(c, POFD_E_ARG, "calcLikeDoubleSingle::setRRange: Data not loaded");
    s.maxsig1 = s.fields[0].sigma_base1;
    s.maxsig2 = s.fields[0].sigma_base2;
    s.mind1 = s.fields[0].mn1;
    s.maxd1 = s.fields[0].mx1;
    s.mind2 = s.fields[0].mn2;
    s.maxd2 = s.fields[0].mx2;
    for (auto& f : s.fields) {
      s.maxsig1 = std::max(s.maxsig1, f.sigma_base1);
      s.maxsig2 = std::max(s.maxsig2, f.sigma_base2);
      s.mind1 = std::min(s.mind1, f.mn1);
      s.maxd1 = std::max(s.maxd1, f.mx1);
      s.mind2 = std::min(s.mind2, f.mn2);
      s.maxd2 = std::max(s.maxd2, f.mx2);
      const double var1 = f.sigma_base1 * f.sigma_base1 + C2->exp_conf1 * C2->exp_conf1;
      const double var2 = f.sigma_base2 * f.sigma_base2 + C2->exp_conf2 * C2->exp_conf2;
      const double sigprod = std::sqrt(var1 * var2);
      f.like_offset = 0.0;
      if (sigprod > 0)
        f.like_offset = -static_cast<double>(f.npix) * (std::log(2.0 * kPi2 * sigprod) + 1.0);
    }
    // getRRangeInternal (:1481-1547)
    const pofd_dbeam& bm = s.bm;
    double r1lo = 0.0, r1hi = 0.0, r2lo = 0.0, r2hi = 0.0, cv;
    if (bm.hassign[0]) {
      cv = mx1 * bm.max1[0];
      if (cv > r1hi) r1hi = cv;
      cv = mx2 * bm.max2[0];
      if (cv > r2hi) r2hi = cv;
    }
    if (bm.hassign[1]) {
      cv = mx1 * bm.max1[1];
      if (cv > r1hi) r1hi = cv;
      cv = -mx2 * bm.max2[1];
      if (cv < r2lo) r2lo = cv;
    }
    if (bm.hassign[2]) {
      cv = -mx1 * bm.max1[2];
      if (cv < r1lo) r1lo = cv;
      cv = mx2 * bm.max2[2];
      if (cv > r2hi) r2hi = cv;
    }
    if (bm.hassign[3]) {
      cv = -mx1 * bm.max1[3];
      if (cv < r1lo) r1lo = cv;
      cv = -mx2 * bm.max2[3];
      if (cv < r2lo) r2lo = cv;
    }
    // setRRange (:496-534)
    const double safetyfac = 1.1, n_zero_pad = 7.5;
    const double sigma1 = std::sqrt(C2->exp_conf1 * C2->exp_conf1 + s.maxsig1 * s.maxsig1);
    const double sigma2 = std::sqrt(C2->exp_conf2 * C2->exp_conf2 + s.maxsig2 * s.maxsig2);
    s.range[0] = safetyfac * r1lo;
    s.range[1] = safetyfac * r1hi + n_zero_pad * sigma1;
    s.range[2] = safetyfac * r2lo;
    s.range[3] = safetyfac * r2hi + n_zero_pad * sigma2;
    if (s.mind1 < s.range[0] && (bm.hassign[2] || bm.hassign[3])) s.range[0] = s.mind1;
    if (s.maxd1 > s.range[1] && (bm.hassign[0] || bm.hassign[1])) s.range[1] = s.maxd1;
    if (s.mind2 < s.range[2] && (bm.hassign[1] || bm.hassign[3])) s.range[2] = s.mind2;
    if (s.maxd2 > s.range[3] && (bm.hassign[0] || bm.hassign[2])) s.range[3] = s.maxd2;
    if ((rc = prepare_set2(c, *C2, s))) return rc;
  }
  return POFD_OK;
}

int pofd2d_set_rrange(pofd_ctx* c, int b, const double* range) {
  Ctx2* C2 = get2(c);
  if (!C2 || !range || b < 0 || static_cast<size_t>(b) >= C2->sets.size())
    return ctx_fail(c, POFD_E_ARG, "pofd2d_set_rrange: bad beam set");
  Set2& s = *C2->sets[b];
  if (std::isnan(s.maxsig1))
    return ctx_fail(c, POFD_E_ARG, "pofd2d_set_rrange: call pofd2d_prepare first");
  for (int k = 0; k < 4; ++k) s.range[k] = range[k];
  return prepare_set2(c, *C2, s);
}

int pofd2d_beamset_info(pofd_ctx* c, int b, double* out) {
  Ctx2* C2 = get2(c);
  if (!C2 || !out || b < 0 || static_cast<size_t>(b) >= C2->sets.size()) return POFD_E_ARG;
  const Set2& s = *C2->sets[b];
  for (int k = 0; k < 4; ++k) out[k] = s.range[k];
  out[4] = s.maxsig1;
  out[5] = s.maxsig2;
  out[6] = s.mind1;
  out[7] = s.maxd1;
  out[8] = s.mind2;
  out[9] = s.maxd2;
  out[10] = static_cast<double>(s.fields.size());
  out[11] = s.fields.empty() ? NAN : s.fields[0].like_offset;
  out[12] = s.prepared ? s.dev.dflux1 : NAN;
  out[13] = s.prepared ? s.dev.dflux2 : NAN;
  return POFD_OK;
}

int pofd2d_loglike_batch_device(pofd_ctx* c, const float* d_theta, uint32_t nwalk,
                                uint32_t stride, double* d_ll, int32_t* d_st, void* stream) {
  Ctx2* C2 = get2(c);
  int rc = check_ready2(c, C2);
  if (rc) return rc;
  if (nwalk == 0) return POFD_OK;
  if (!d_theta || !d_ll || !d_st)
    return ctx_fail(c, POFD_E_ARG, "pofd2d_loglike_batch_device: null buffer");
  const LNHost& m = C2->model;
  if (stride < static_cast<uint32_t>(m.nk + m.ns + m.no + 2))
    return ctx_fail(c, POFD_E_ARG, "calcLikeDouble::getLogLike: Not enough elements in params");
  for (size_t b = 0; b < C2->sets.size(); ++b) {
    Set2& s = *C2->sets[b];
    if ((rc = ensure_work2(c, *C2, s, nwalk))) return rc;
    s.dev.spectral_cut = pofd2::ctx_spectral_cut(c);
    hipEvent_t* ev = nullptr;
    if ((rc = pofd2::ctx_stage_events(c, &ev))) return rc;
    s.dev.bonus = (b == 0) ? 1 : 0;  // the counts integrals once per batch
    rc = pofd2::launch_loglike2d(s.dev, s.wk, s.gmain, s.g00, s.gx, s.gy, d_theta, stride, d_ll,
                                 d_st, b > 0 ? 1 : 0, stream, ev);
    if (rc)
      return ctx_fail(c, rc, std::string("pofd2d_loglike_batch_device: launch failed: ") +
                                 hipGetErrorString(hipGetLastError()));
  }
  if (C2->priors.flags || C2->priors.reg_alpha > 0.0) {
    Set2& s0 = *C2->sets[0];
    rc = pofd2::launch_prior2d(s0.dev, s0.wk, C2->priors, d_theta, stride, nwalk, d_ll, d_st,
                               stream);
    if (rc) return ctx_fail(c, rc, "pofd2d_loglike_batch_device: prior launch failed");
  }
  C2->last_nwalk = nwalk;
  return POFD_OK;
}

int pofd2d_set_priors(pofd_ctx* c, int flags, const double* sigma_width, const double* cfirb,
                      const double* poisson, double reg_alpha) {
  Ctx2* C2 = get2(c);
  if (!C2) return POFD_E_ARG;
  if (flags & ~63) return ctx_fail(c, POFD_E_ARG, "pofd2d_set_priors: unknown prior flag");
  using namespace pofd2;
  Prior2Dev p{};
  p.flags = flags;
  p.reg_alpha = reg_alpha;
  for (int b = 0; b < 2; ++b) {
    if (flags & (kP2Sigma1 << b)) {
      if (!sigma_width || !(sigma_width[b] > 0.0))
        return ctx_fail(c, POFD_E_ARG, "pofd2d_set_priors: invalid sigma prior width");
      p.sigma_width[b] = sigma_width[b];
    }
    if (flags & (kP2Cfirb1 << b)) {
      if (!cfirb || !(cfirb[2 * b + 1] > 0.0))
        return ctx_fail(c, POFD_E_ARG, "pofd2d_set_priors: invalid CFIRB prior");
      p.cfirb_mean[b] = cfirb[2 * b];
      p.cfirb_sigma[b] = cfirb[2 * b + 1];
    }
    if (flags & (kP2Poisson1 << b)) {
      if (!poisson || !(poisson[2 * b + 1] > 0.0))
        return ctx_fail(c, POFD_E_ARG, "pofd2d_set_priors: invalid Poisson prior");
      p.poisson_mean[b] = poisson[2 * b];
      p.poisson_sigma[b] = poisson[2 * b + 1];
    }
  }
  if (reg_alpha < 0.0) return ctx_fail(c, POFD_E_ARG, "pofd2d_set_priors: negative alpha");
  C2->priors = p;
  return POFD_OK;
}

int pofd2d_loglike_batch(pofd_ctx* c, const float* theta, uint32_t nwalk, uint32_t stride,
                         double* ll, int32_t* st) {
  Ctx2* C2 = get2(c);
  int rc = check_ready2(c, C2);
  if (rc) return rc;
  if (nwalk == 0) return POFD_OK;
  if (!theta || !ll || !st) return ctx_fail(c, POFD_E_ARG, "pofd2d_loglike_batch: null buffer");
  const size_t tb = static_cast<size_t>(nwalk) * stride * sizeof(float);
  if ((rc = alloc(c, C2->dtheta, tb))) return rc;
  if ((rc = alloc(c, C2->dll, nwalk * sizeof(double)))) return rc;
  if ((rc = alloc(c, C2->dst, nwalk * sizeof(int32_t)))) return rc;
  hipError_t e = hipMemcpy(C2->dtheta->p, theta, tb, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hipfail(c, e, "hipMemcpy(theta)");
  rc = pofd2d_loglike_batch_device(c, static_cast<const float*>(C2->dtheta->p), nwalk, stride,
                                   static_cast<double*>(C2->dll->p),
                                   static_cast<int32_t*>(C2->dst->p), nullptr);
  if (rc) return rc;
  e = hipMemcpy(ll, C2->dll->p, nwalk * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(st, C2->dst->p, nwalk * sizeof(int32_t), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hipfail(c, e, "hipMemcpy(results)");
  return POFD_OK;
}

// info[32]: the per-(walker, dataset) record: status, split1, split2,
// sg1pos, sg2pos, sigma1, sigma2, integral, 1/integral, minflux1, minflux2,
// dataset logL (before offset), then mn1, mn2, var1, var2, r0, K1, K2
// the counts integrals of the last batch (beam set 0's workspace; 8 per
// batch row: NS, S1, S2, S1^2, S2^2 per area), for the sampler's bonus
// parameters; null before any batch
const double* pofd2d_bonus_device(pofd_ctx* c) {
  Ctx2* C2 = get2(c);
  if (!C2 || C2->sets.empty() || !C2->sets[0]->wk.bonus) return nullptr;
  return C2->sets[0]->wk.bonus;
}

int pofd2d_get_bonus(pofd_ctx* c, uint32_t nwalk, double* out) {
  Ctx2* C2 = get2(c);
  if (!C2 || !out || C2->sets.empty() || !C2->sets[0]->wbuf || nwalk > C2->last_nwalk)
    return ctx_fail(c, POFD_E_ARG, "pofd2d_get_bonus: no such walkers in the last batch");
  const Set2& s = *C2->sets[0];
  std::vector<double> tmp(static_cast<size_t>(nwalk) * 8);
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess)
    e = hipMemcpy(tmp.data(), s.wk.bonus, tmp.size() * sizeof(double), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hipfail(c, e, "hipMemcpy(bonus)");
  for (uint32_t w = 0; w < nwalk; ++w)
    for (int k = 0; k < 5; ++k) out[5 * w + k] = tmp[8 * static_cast<size_t>(w) + k];
  return POFD_OK;
}

int pofd2d_get_debug(pofd_ctx* c, uint32_t walker, int b, int f, double* info) {
  Ctx2* C2 = get2(c);
  if (!C2 || !info || b < 0 || static_cast<size_t>(b) >= C2->sets.size())
    return ctx_fail(c, POFD_E_ARG, "pofd2d_get_debug: bad beam set");
  Set2& s = *C2->sets[b];
  if (!s.wbuf || walker >= s.wcap || walker >= C2->last_nwalk || f < 0 ||
      static_cast<size_t>(f) >= s.fields.size())
    return ctx_fail(c, POFD_E_ARG, "pofd2d_get_debug: no such walker/dataset in the last batch");
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hipfail(c, e, "hipDeviceSynchronize");
  const size_t wf = static_cast<size_t>(walker) * s.fields.size() + f;
  e = hipMemcpy(info, s.wk.dbg + wf * pofd2::kDbg2, pofd2::kDbg2 * sizeof(double),
                hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hipfail(c, e, "hipMemcpy(debug)");
  double st[pofd2::kStats], mh[8];
  e = hipMemcpy(st, s.wk.stats + static_cast<size_t>(walker) * pofd2::kStats, sizeof(st),
                hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    e = hipMemcpy(mh, s.wk.model + static_cast<size_t>(walker) * pofd2::kModelStride, sizeof(mh),
                  hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hipfail(c, e, "hipMemcpy(stats)");
  info[pofd2::D2_MN1] = st[pofd2::ST2_MN1];
  info[pofd2::D2_MN2] = st[pofd2::ST2_MN2];
  info[pofd2::D2_VAR1] = st[pofd2::ST2_VAR1];
  info[pofd2::D2_VAR2] = st[pofd2::ST2_VAR2];
  info[pofd2::D2_R0] = st[pofd2::ST2_R0];
  info[pofd2::D2_K1] = mh[3];
  info[pofd2::D2_K2] = mh[4];
  if (mh[0] != POFD_OK) info[pofd2::D2_STATUS] = mh[0];
  return POFD_OK;
}

// the rotated, normalised P(D) of the last batch as PDFactoryDouble::getPD
// leaves PDDouble::pd_ (log2 when is_log), info[8] = minflux1, minflux2,
// dflux1, dflux2, split1, split2, sg1pos, sg2pos
int pofd2d_get_pd(pofd_ctx* c, uint32_t walker, int b, int f, int is_log, double* pd_out,
                  double* info) {
  double rec[pofd2::kDbg2];
  int rc = pofd2d_get_debug(c, walker, b, f, rec);
  if (rc) return rc;
  if (rec[pofd2::D2_STATUS] != POFD_OK)
    return ctx_fail(c, static_cast<int>(rec[pofd2::D2_STATUS]),
                    "pofd2d_get_pd: the P(D) of this walker failed");
  Ctx2* C2 = get2(c);
  Set2& s = *C2->sets[b];
  const size_t n = C2->n;
  std::vector<double> P(n * n);
  const size_t wf = static_cast<size_t>(walker) * s.fields.size() + f;
  hipError_t e = hipMemcpy(P.data(), s.wk.P + wf * s.wk.P_stride, n * n * sizeof(double),
                           hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hipfail(c, e, "hipMemcpy(P)");
  const size_t s1 = static_cast<size_t>(rec[pofd2::D2_SPLIT1]);
  const size_t s2 = static_cast<size_t>(rec[pofd2::D2_SPLIT2]);
  const double itot = rec[pofd2::D2_ITOT];
  if (pd_out)
    for (size_t i = 0; i < n; ++i)
      for (size_t j = 0; j < n; ++j) {
        const double v = P[((i + s1) % n) * n + (j + s2) % n] * itot;
        pd_out[i * n + j] = is_log ? ((v > 0.0) ? std::log2(v) : -30.0) : v;
      }
  if (info) {
    info[0] = rec[pofd2::D2_MINFLUX1];
    info[1] = rec[pofd2::D2_MINFLUX2];
    info[2] = s.dev.dflux1;
    info[3] = s.dev.dflux2;
    info[4] = rec[pofd2::D2_SPLIT1];
    info[5] = rec[pofd2::D2_SPLIT2];
    info[6] = rec[pofd2::D2_SG1POS];
    info[7] = rec[pofd2::D2_SG2POS];
  }
  return POFD_OK;
}

}  // extern "C"
