/*
 * pofd_oracle2d.c -- CPU restatement of the reference 2D P(D) likelihood.
 * TEST INFRASTRUCTURE ONLY (see pofd_oracle2d.h).  Every function cites the
 * reference file:line it follows.  Plain double arithmetic, built with
 * -ffp-contract=off like the rest of the oracle.
 */
#include "pofd_oracle2d.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC2_PI 3.14159265358979323846
static const double orc2_two_pi = 2.0 * ORC2_PI;
static const double orc2_n_zero_pad = 7.5;    /* pofd_mcmc::n_zero_pad */
static const double orc2_smalllogval = -30.;  /* pofd_mcmc::smalllogval */

static int orc2_fail(orc_err *e, int code, const char *cls, const char *method,
                     const char *fmt, ...) {
  if (e) {
    e->code = code;
    snprintf(e->cls, sizeof e->cls, "%s", cls);
    snprintf(e->method, sizeof e->method, "%s", method);
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(e->msg, sizeof e->msg, fmt, ap);
    va_end(ap);
  }
  return code;
}

/* ===================================================================== */
/* doublebeam (src/doublebeam.cc)                                        */
/* ===================================================================== */

orc_dbeam *orc_dbeam_new(void) {
  orc_dbeam *b = (orc_dbeam *)calloc(1, sizeof(orc_dbeam));
  return b;
}

static void orc_dbeam_clear(orc_dbeam *b) {
  for (int s = 0; s < 4; ++s) {
    free(b->pix1[s]);
    free(b->pix2[s]);
    free(b->ipix1[s]);
    free(b->ipix2[s]);
    free(b->logratio[s]);
    free(b->binw[s]);
    free(b->binv1[s]);
    free(b->binv2[s]);
    free(b->binlogratio[s]);
  }
  memset(b, 0, sizeof *b);
}

void orc_dbeam_free(orc_dbeam *b) {
  if (!b) return;
  orc_dbeam_clear(b);
  free(b);
}

/* doublebeam::setBeams, src/doublebeam.cc:331-460; the log ratios of
   getLogRatio (:722-748) are formed here instead of on first use */
int orc_dbeam_set(orc_dbeam *bm, const double *beam1, const double *beam2,
                  size_t n, double pixsize, double minval, orc_err *err) {
  orc_dbeam_clear(bm);
  if (n == 0) return orc2_fail(err, ORC_E_ARG, "doublebeam", "setBeams", "Invalid (zero) n");
  if (minval < 0.0)
    return orc2_fail(err, ORC_E_ARG, "doublebeam", "setBeams", "Invalid (negative) MINVAL");
  bm->minval = minval;
  unsigned *component = (unsigned *)malloc(sizeof(unsigned) * n);
  for (size_t i = 0; i < n; ++i) {
    double val1 = beam1[i], val2 = beam2[i];
    double aval1 = fabs(val1), aval2 = fabs(val2);
    if ((aval1 > minval) && (aval2 > minval)) {
      unsigned sgn;
      if (val1 > 0) sgn = (val2 > 0) ? 0 : 1;
      else sgn = (val2 > 0) ? 2 : 3;
      component[i] = sgn;
      bm->npix[sgn] += 1;
    } else {
      component[i] = 4;
    }
  }
  unsigned totpix = bm->npix[0] + bm->npix[1] + bm->npix[2] + bm->npix[3];
  if (totpix == 0) {
    free(component);
    return orc2_fail(err, ORC_E_ARG, "doublebeam", "setBeams", "No acceptable values found");
  }
  for (int s = 0; s < 4; ++s) {
    unsigned cn = bm->npix[s];
    if (cn == 0) continue;
    bm->hassign[s] = 1;
    double *p1 = bm->pix1[s] = (double *)malloc(sizeof(double) * cn);
    double *p2 = bm->pix2[s] = (double *)malloc(sizeof(double) * cn);
    bm->ipix1[s] = (double *)malloc(sizeof(double) * cn);
    bm->ipix2[s] = (double *)malloc(sizeof(double) * cn);
    unsigned idx = 0;
    for (size_t j = 0; j < n; ++j)
      if (component[j] == (unsigned)s) {
        p1[idx] = fabs(beam1[j]);
        p2[idx] = fabs(beam2[j]);
        ++idx;
      }
    double val = p1[0];
    bm->tot1[s] = val;
    bm->totsq1[s] = val * val;
    for (unsigned j = 1; j < cn; ++j) {
      val = p1[j];
      bm->tot1[s] += val;
      bm->totsq1[s] += val * val;
    }
    val = p2[0];
    bm->tot2[s] = val;
    bm->totsq2[s] = val * val;
    for (unsigned j = 1; j < cn; ++j) {
      val = p2[j];
      bm->tot2[s] += val;
      bm->totsq2[s] += val * val;
    }
    for (unsigned j = 0; j < cn; ++j) bm->ipix1[s][j] = 1.0 / p1[j];
    for (unsigned j = 0; j < cn; ++j) bm->ipix2[s][j] = 1.0 / p2[j];
  }
  free(component);
  for (int s = 0; s < 4; ++s) {
    if (!bm->hassign[s]) {
      bm->min1[s] = bm->max1[s] = bm->min2[s] = bm->max2[s] = NAN;
      continue;
    }
    unsigned cn = bm->npix[s];
    const double *p1 = bm->pix1[s], *p2 = bm->pix2[s];
    double mn = p1[0], mx = p1[0];
    for (unsigned j = 1; j < cn; ++j) {
      double v = p1[j];
      if (v > mx) mx = v;
      else if (v < mn) mn = v;
    }
    bm->min1[s] = mn;
    bm->max1[s] = mx;
    mn = mx = p2[0];
    for (unsigned j = 1; j < cn; ++j) {
      double v = p2[j];
      if (v > mx) mx = v;
      else if (v < mn) mn = v;
    }
    bm->min2[s] = mn;
    bm->max2[s] = mx;
    /* getLogRatio: log(beam1 * (1/beam2)) */
    bm->logratio[s] = (double *)malloc(sizeof(double) * cn);
    for (unsigned j = 0; j < cn; ++j) bm->logratio[s][j] = log(p1[j] * bm->ipix2[s][j]);
  }
  if (pixsize <= 0.0)
    return orc2_fail(err, ORC_E_ARG, "doublebeam", "setBeams", "Invalid (non-positive) pixel size");
  bm->pixsize = pixsize;
  return ORC_OK;
}

/* doublebeam::makeHistogram, src/doublebeam.cc:467-584 (histothresh 30,
   include/doublebeam.h:36); getBinLogRatio :750-776 formed here */
int orc_dbeam_make_histogram(orc_dbeam *bm, uint32_t NBINS, orc_err *err) {
  int has = 0;
  for (int s = 0; s < 4; ++s) has |= bm->hassign[s];
  if (!has)
    return orc2_fail(err, ORC_E_ARG, "doublebeam", "makeHistogram", "Trying to histogram empty beam");
  if (NBINS == 0)
    return orc2_fail(err, ORC_E_ARG, "doublebeam", "makeHistogram", "Invalid (zero) NBINS");
  bm->nbins = NBINS;
  for (int s = 0; s < 4; ++s) {
    bm->ishist[s] = 0;
    bm->nhist[s] = 0;
    free(bm->binw[s]);
    free(bm->binv1[s]);
    free(bm->binv2[s]);
    free(bm->binlogratio[s]);
    bm->binw[s] = bm->binv1[s] = bm->binv2[s] = bm->binlogratio[s] = NULL;
  }
  const unsigned histothresh = 30;
  const size_t nbins2 = (size_t)NBINS * NBINS;
  unsigned *tmpwt = (unsigned *)malloc(sizeof(unsigned) * nbins2);
  double *tmph1 = (double *)malloc(sizeof(double) * nbins2);
  double *tmph2 = (double *)malloc(sizeof(double) * nbins2);
  const double log2outscale = 0.0014419741739063218;
  const double dnbins = (double)NBINS;
  for (int s = 0; s < 4; ++s) {
    if (!(bm->hassign[s] && bm->npix[s] >= histothresh)) continue;
    double minbinval1 = log2(bm->min1[s]) - log2outscale;
    double maxbinval1 = log2(bm->max1[s]) + log2outscale;
    double ihiststep1 = dnbins / (maxbinval1 - minbinval1);
    double minbinval2 = log2(bm->min2[s]) - log2outscale;
    double maxbinval2 = log2(bm->max2[s]) + log2outscale;
    double ihiststep2 = dnbins / (maxbinval2 - minbinval2);
    memset(tmpwt, 0, sizeof(unsigned) * nbins2);
    memset(tmph1, 0, sizeof(double) * nbins2);
    memset(tmph2, 0, sizeof(double) * nbins2);
    const double *p1 = bm->pix1[s], *p2 = bm->pix2[s];
    for (unsigned i = 0; i < bm->npix[s]; ++i) {
      double val1 = p1[i], val2 = p2[i];
      unsigned idx1 = (unsigned)((log2(val1) - minbinval1) * ihiststep1);
      unsigned idx2 = (unsigned)((log2(val2) - minbinval2) * ihiststep2);
      size_t tot = (size_t)idx1 * NBINS + idx2;
      tmpwt[tot] += 1;
      tmph1[tot] += val1;
      tmph2[tot] += val2;
    }
    unsigned nnz = 0;
    for (size_t i = 0; i < nbins2; ++i)
      if (tmpwt[i] > 0) ++nnz;
    bm->binw[s] = (double *)malloc(sizeof(double) * nnz);
    bm->binv1[s] = (double *)malloc(sizeof(double) * nnz);
    bm->binv2[s] = (double *)malloc(sizeof(double) * nnz);
    bm->binlogratio[s] = (double *)malloc(sizeof(double) * nnz);
    unsigned t = 0;
    for (size_t i = 0; i < nbins2; ++i) {
      unsigned u = tmpwt[i];
      if (u > 0) {
        double wn = (double)u;
        bm->binw[s][t] = wn;
        bm->binv1[s][t] = wn / tmph1[i];
        bm->binv2[s][t] = wn / tmph2[i];
        ++t;
      }
    }
    for (unsigned i = 0; i < nnz; ++i)
      bm->binlogratio[s][i] = log(bm->binv2[s][i] / bm->binv1[s][i]);
    bm->ishist[s] = 1;
    bm->nhist[s] = nnz;
  }
  free(tmpwt);
  free(tmph1);
  free(tmph2);
  return ORC_OK;
}

void orc_dbeam_info(const orc_dbeam *bm, double *out) {
  for (int s = 0; s < 4; ++s) {
    double *o = out + 12 * s;
    o[0] = bm->hassign[s];
    o[1] = bm->npix[s];
    o[2] = bm->ishist[s];
    o[3] = bm->nhist[s];
    o[4] = bm->hassign[s] ? bm->tot1[s] : 0.0;
    o[5] = bm->hassign[s] ? bm->tot2[s] : 0.0;
    o[6] = bm->hassign[s] ? bm->totsq1[s] : 0.0;
    o[7] = bm->hassign[s] ? bm->totsq2[s] : 0.0;
    o[8] = bm->min1[s];
    o[9] = bm->max1[s];
    o[10] = bm->min2[s];
    o[11] = bm->max2[s];
  }
  out[48] = bm->pixsize;
  out[49] = bm->minval;
  out[50] = bm->nbins;
}

const double *orc_dbeam_array(const orc_dbeam *bm, int s, int which,
                              uint32_t *n) {
  if (s < 0 || s > 3) {
    *n = 0;
    return NULL;
  }
  *n = (which <= 4) ? bm->npix[s] : bm->nhist[s];
  switch (which) {
    case 0: return bm->pix1[s];
    case 1: return bm->pix2[s];
    case 2: return bm->ipix1[s];
    case 3: return bm->ipix2[s];
    case 4: return bm->logratio[s];
    case 5: return bm->binw[s];
    case 6: return bm->binv1[s];
    case 7: return bm->binv2[s];
    case 8: return bm->binlogratio[s];
  }
  *n = 0;
  return NULL;
}

/* ===================================================================== */
/* numberCountsDoubleLogNormal (src/numberCountsDoubleLogNormal.cc)      */
/* ===================================================================== */

void orc_ln_init(orc_lnmodel *m) {
  memset(m, 0, sizeof *m);
  m->min_sigma = 0.0;
}

/* setKnotPositions / setSigmaPositions / setOffsetPositions
   (:632-656, :744-767, :873-895) */
int orc_ln_set_positions(orc_lnmodel *m, unsigned nk, const double *kp,
                         unsigned ns, const double *sp, unsigned no,
                         const double *op, orc_err *err) {
  if (nk <= 2 || nk > ORC_MAXK2 || ns < 1 || ns > ORC_MAXK2 || no < 1 || no > ORC_MAXK2)
    return orc2_fail(err, ORC_E_ARG, "numberCountsDoubleLogNormal", "setPositions",
                     "need > 2 knots and >= 1 sigma and offset knots");
  for (unsigned i = 0; i < nk; ++i)
    if (kp[i] <= 0.0)
      return orc2_fail(err, ORC_E_ARG, "numberCountsDoubleLogNormal", "setKnotPositions",
                       "Non-positive knot positions not allowed");
  for (unsigned i = 1; i < nk; ++i)
    if (kp[i] <= kp[i - 1])
      return orc2_fail(err, ORC_E_ARG, "numberCountsDoubleLogNormal", "setKnotPositions",
                       "Knot positions not monotonically increasing");
  for (unsigned i = 0; i < ns; ++i)
    if (sp[i] <= 0.0)
      return orc2_fail(err, ORC_E_ARG, "numberCountsDoubleLogNormal", "setSigmaPositions",
                       "Non-positive sigma knot positions not allowed");
  for (unsigned i = 0; i < no; ++i)
    if (op[i] <= 0.0)
      return orc2_fail(err, ORC_E_ARG, "numberCountsDoubleLogNormal", "setOffsetPositions",
                       "Non-positive offset knot positions not allowed");
  m->nknots = nk;
  m->nsigma = ns;
  m->noffset = no;
  for (unsigned i = 0; i < nk; ++i) {
    m->knots[i] = kp[i];
    m->logknots[i] = log2(kp[i]);
  }
  for (unsigned i = 0; i < ns; ++i) m->sigmaknots[i] = sp[i];
  for (unsigned i = 0; i < no; ++i) m->offsetknots[i] = op[i];
  m->pos_loaded = 1;
  m->vals_loaded = 0;
  m->knots_valid = m->sigmas_valid = m->offsets_valid = 0;
  return ORC_OK;
}

/* setMinSigma (:792-803) */
int orc_ln_set_min_sigma(orc_lnmodel *m, double mins, orc_err *err) {
  if (mins < 0.0)
    return orc2_fail(err, ORC_E_ARG, "numberCountsDoubleLogNormal", "setMinSigma",
                     "Invalid (negative) min sigma");
  m->min_sigma = mins;
  return ORC_OK;
}

/* setParams (:1010-1052) with checkKnotsValid / checkSigmasValid /
   checkOffsetsValid (:1060-1119) */
int orc_ln_set_params(orc_lnmodel *m, unsigned npar, const float *F, orc_err *err) {
  const unsigned nk = m->nknots, ns = m->nsigma, no = m->noffset;
  if (nk + ns + no > npar)
    return orc2_fail(err, ORC_E_ARG, "numberCountsDoubleLogNormal", "setParams",
                     "Not enough parameters present to set");
  if (!m->pos_loaded)
    return orc2_fail(err, ORC_E_ARG, "numberCountsDoubleLogNormal", "setParams",
                     "Knot positions not set");
  const double logfac = log2(10.0);
  for (unsigned i = 0; i < nk; ++i) m->logknotvals[i] = logfac * (double)F[i];
  orc_cspline_init(nk, m->logknots, m->logknotvals, m->kc);
  for (unsigned i = 0; i < ns; ++i) m->sigmavals[i] = (double)F[nk + i];
  if (ns > 2) orc_cspline_init(ns, m->sigmaknots, m->sigmavals, m->sc);
  for (unsigned i = 0; i < no; ++i) m->offsetvals[i] = (double)F[nk + ns + i];
  if (no > 2) orc_cspline_init(no, m->offsetknots, m->offsetvals, m->oc);
  m->vals_loaded = 1;
  int ok = 1;
  for (unsigned i = 0; i < nk; ++i)
    if (isnan(m->knots[i]) || isinf(m->knots[i])) ok = 0;
  if (m->knots[0] <= 0.0) ok = 0;
  for (unsigned i = 1; i < nk; ++i)
    if (m->knots[i] <= m->knots[i - 1]) ok = 0;
  for (unsigned i = 0; i < nk; ++i)
    if (isnan(m->logknotvals[i])) ok = 0;
  m->knots_valid = ok;
  ok = (m->min_sigma >= 0.0);
  for (unsigned i = 0; i < ns; ++i)
    if (isnan(m->sigmaknots[i]) || isinf(m->sigmaknots[i])) ok = 0;
  if (m->sigmaknots[0] <= 0.0) ok = 0;
  for (unsigned i = 1; i < ns; ++i)
    if (m->sigmaknots[i] <= m->sigmaknots[i - 1]) ok = 0;
  for (unsigned i = 0; i < ns; ++i)
    if (isnan(m->sigmavals[i]) || isinf(m->sigmavals[i])) ok = 0;
  m->sigmas_valid = ok;
  ok = 1;
  for (unsigned i = 0; i < no; ++i)
    if (isnan(m->offsetknots[i]) || isinf(m->offsetknots[i])) ok = 0;
  if (m->offsetknots[0] <= 0.0) ok = 0;
  for (unsigned i = 1; i < no; ++i)
    if (m->offsetknots[i] <= m->offsetknots[i - 1]) ok = 0;
  for (unsigned i = 0; i < no; ++i)
    if (isnan(m->offsetvals[i])) ok = 0;
  m->offsets_valid = ok;
  return ORC_OK;
}

int orc_ln_is_valid(const orc_lnmodel *m) {
  return m->vals_loaded && m->knots_valid && m->sigmas_valid && m->offsets_valid;
}

/* gsl_interp_linear evaluation (GSL interp/linear.c linear_eval) */
static double orc2_linear_eval(const double *x, const double *y, size_t n,
                               double xv) {
  size_t i = orc_bsearch(x, xv, 0, n - 1);
  double xlo = x[i], xhi = x[i + 1], ylo = y[i], yhi = y[i + 1];
  double dx = xhi - xlo;
  if (dx > 0.0) return ylo + (xv - xlo) / dx * (yhi - ylo);
  return 0.0;
}

/* getSigmaInner (:1172-1180) */
double orc_ln_sigma(const orc_lnmodel *m, double f1) {
  const unsigned ns = m->nsigma;
  if (ns == 1) return m->sigmavals[0];
  if (f1 <= m->sigmaknots[0]) return m->sigmavals[0];
  if (f1 >= m->sigmaknots[ns - 1]) return m->sigmavals[ns - 1];
  double val = (ns == 2) ? orc2_linear_eval(m->sigmaknots, m->sigmavals, ns, f1)
                         : orc_cspline_eval(ns, m->sigmaknots, m->sigmavals, m->sc, f1);
  return (val < m->min_sigma) ? m->min_sigma : val;
}

/* getOffsetInner (:1187-1193) */
double orc_ln_offset(const orc_lnmodel *m, double f1) {
  const unsigned no = m->noffset;
  if (no == 1) return m->offsetvals[0];
  if (f1 <= m->offsetknots[0]) return m->offsetvals[0];
  if (f1 >= m->offsetknots[no - 1]) return m->offsetvals[no - 1];
  return (no == 2) ? orc2_linear_eval(m->offsetknots, m->offsetvals, no, f1)
                   : orc_cspline_eval(no, m->offsetknots, m->offsetvals, m->oc, f1);
}

/* getNumberCountsInner (:1230-1254) */
static double orc2_counts_inner(const orc_lnmodel *m, double f1, double f2) {
  const double normfac = 1.0 / sqrt(2 * ORC2_PI);
  const double prefac = -0.5 * (1.0 / log(2.0)); /* -0.5 * ilog2toe */
  if (f1 < m->knots[0] || f1 >= m->knots[m->nknots - 1] || f2 <= 0.0) return 0.0;
  double cnts = exp2(orc_cspline_eval(m->nknots, m->logknots, m->logknotvals, m->kc, log2(f1)));
  double if1 = 1.0 / f1;
  double isigma = 1.0 / orc_ln_sigma(m, f1);
  double tfac = (log(f2 * if1) - orc_ln_offset(m, f1)) * isigma;
  cnts *= normfac * isigma * exp2(prefac * tfac * tfac) / f2;
  return cnts;
}

double orc_ln_counts(const orc_lnmodel *m, double f1, double f2) {
  if (!orc_ln_is_valid(m)) return NAN;
  if (isnan(f1) || isinf(f1) || isnan(f2) || isinf(f2)) return NAN;
  return orc2_counts_inner(m, f1, f2);
}

/* getBand1NumberCounts (:1272-1288) */
double orc_ln_band1_counts(const orc_lnmodel *m, double f1) {
  if (!orc_ln_is_valid(m)) return NAN;
  if (f1 < m->knots[0] || f1 >= m->knots[m->nknots - 1]) return 0.0;
  return exp2(orc_cspline_eval(m->nknots, m->logknots, m->logknotvals, m->kc, log2(f1)));
}

static int orc2_cmp_double(const void *a, const void *b) {
  const double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

/* numberCountsDoubleLogNormal::splineInt(alpha, beta),
   src/numberCountsDoubleLogNormal.cc:1549-1610 with the integrand
   evalPowfNDoubleLogNormal (:2424-2519): int S1^(alpha+beta)
   exp(beta mu(S1) + beta^2/2 sigma(S1)^2) n1(S1) dS1 over [minknot, maxknot)
   (sigma floored at min_sigma).  The reference integrates with GSL QAG-41 to
   relative 1e-5; here the integral is taken to full double precision as
   32-point Gauss-Legendre sums in log2 S1 between consecutive band-1, sigma
   and offset knots, where the integrand is analytic. */
double orc_ln_splineint(const orc_lnmodel *m, double alpha, double beta) {
  if (m->nknots < 2 || !orc_ln_is_valid(m)) return NAN;
  enum { NG = 32 };
  double gx[NG], gw[NG];
  orc_gl_nodes(NG, gx, gw);
  double bp[3 * ORC_MAXK2];
  unsigned nb = 0;
  const double lo = m->logknots[0], hi = m->logknots[m->nknots - 1];
  for (unsigned i = 0; i < m->nknots; ++i) bp[nb++] = m->logknots[i];
  for (unsigned i = 0; i < m->nsigma; ++i) {
    const double v = log2(m->sigmaknots[i]);
    if (v > lo && v < hi) bp[nb++] = v;
  }
  for (unsigned i = 0; i < m->noffset; ++i) {
    const double v = log2(m->offsetknots[i]);
    if (v > lo && v < hi) bp[nb++] = v;
  }
  qsort(bp, nb, sizeof(double), orc2_cmp_double);
  const double power = alpha + beta, c1 = beta, c2 = 0.5 * beta * beta;
  long double acc = 0.0L;
  for (unsigned k = 0; k + 1 < nb; ++k) {
    const double v0 = bp[k], v1 = bp[k + 1];
    if (!(v1 > v0)) continue;
    const double h = 0.5 * (v1 - v0), c = 0.5 * (v1 + v0);
    for (unsigned g = 0; g < NG; ++g) {
      const double v = c + h * gx[g];
      const double s1 = exp2(v);
      double e = 0.0;
      if (c1 != 0.0) e += c1 * orc_ln_offset(m, s1);
      if (c2 != 0.0) {
        double sg = orc_ln_sigma(m, s1);
        if (sg < m->min_sigma) sg = m->min_sigma;
        e += c2 * sg * sg;
      }
      const double lc = orc_cspline_eval(m->nknots, m->logknots, m->logknotvals, m->kc, v);
      acc += (long double)(gw[g] * h) *
             (long double)(exp2(lc + (power + 1.0) * v) * exp(e));
    }
  }
  return (double)(acc * 0.693147180559945309417232121458L);
}

/* the log-normal density minus the target (static lognorm, :2532-2546) */
typedef struct {
  double mu, sig2, isig, isig2, target;
} orc2_lnp;

static double orc2_lognorm(double u, const orc2_lnp *p) {
  const double prefac = 1.0 / sqrt(2 * ORC2_PI);
  double eu = exp(-u);
  double arg = u - p->mu;
  double L = prefac * exp(-0.5 * arg * arg * p->isig2) * p->isig * eu;
  return L - p->target;
}

/* GSL Brent-Dekker root bracketing (roots/brent.c): state and one
   iteration, restated from the published algorithm */
typedef struct {
  double a, b, c, d, e, fa, fb, fc;
} orc2_brent;

static int orc2_brent_init(orc2_brent *s, const orc2_lnp *p, double lo, double hi,
                           double *root) {
  *root = 0.5 * (lo + hi);
  double flo = orc2_lognorm(lo, p), fhi = orc2_lognorm(hi, p);
  s->a = lo;
  s->fa = flo;
  s->b = hi;
  s->fb = fhi;
  s->c = hi;
  s->fc = fhi;
  s->d = hi - lo;
  s->e = hi - lo;
  if ((flo < 0.0 && fhi < 0.0) || (flo > 0.0 && fhi > 0.0)) return -1;
  return 0;
}

static void orc2_brent_iterate(orc2_brent *s, const orc2_lnp *p, double *root,
                               double *xlo, double *xhi) {
  double a = s->a, b = s->b, c = s->c, d = s->d, e = s->e;
  double fa = s->fa, fb = s->fb, fc = s->fc;
  int ac_equal = 0;
  if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) {
    ac_equal = 1;
    c = a;
    fc = fa;
    d = b - a;
    e = b - a;
  }
  if (fabs(fc) < fabs(fb)) {
    ac_equal = 1;
    a = b;
    b = c;
    c = a;
    fa = fb;
    fb = fc;
    fc = fa;
  }
  double tol = 0.5 * DBL_EPSILON * fabs(b);
  double m = 0.5 * (c - b);
  if (fb == 0) {
    *root = b;
    *xlo = b;
    *xhi = b;
    return;
  }
  if (fabs(m) <= tol) {
    *root = b;
    if (b < c) {
      *xlo = b;
      *xhi = c;
    } else {
      *xlo = c;
      *xhi = b;
    }
    return;
  }
  if (fabs(e) < tol || fabs(fa) <= fabs(fb)) {
    d = m;
    e = m;
  } else {
    double pp, q, r, ss = fb / fa;
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
    double lim1 = 3 * m * q - fabs(tol * q), lim2 = fabs(e * q);
    if (2 * pp < ((lim1 < lim2) ? lim1 : lim2)) {
      e = d;
      d = pp / q;
    } else {
      d = m;
      e = m;
    }
  }
  a = b;
  fa = fb;
  if (fabs(d) > tol)
    b += d;
  else
    b += (m > 0 ? +tol : -tol);
  fb = orc2_lognorm(b, p);
  s->a = a;
  s->b = b;
  s->c = c;
  s->d = d;
  s->e = e;
  s->fa = fa;
  s->fb = fb;
  s->fc = fc;
  *root = b;
  if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) c = a;
  if (b < c) {
    *xlo = b;
    *xhi = c;
  } else {
    *xlo = c;
    *xhi = b;
  }
}

/* gsl_root_test_interval with epsabs = 0 */
static int orc2_test_interval(double lo, double hi, double epsrel) {
  double alo = fabs(lo), ahi = fabs(hi);
  double min_abs = ((lo > 0.0 && hi > 0.0) || (lo < 0.0 && hi < 0.0)) ? ((alo < ahi) ? alo : ahi) : 0.0;
  double tolerance = epsrel * min_abs;
  return fabs(hi - lo) < tolerance;
}

/* logNormalSolver (:1389-1478) */
static int orc2_lognormal_solver(double mu, double sig, double nsig, double *out,
                                 orc_err *err) {
  const unsigned max_bracket_iters = 20;
  const double prefac = 1.0 / sqrt(orc2_two_pi);
  const double tol = 1e-3;
  if (sig <= 0.0)
    return orc2_fail(err, ORC_E_ARG, "numberCountsDoubleLogNormal", "logNormalSolver",
                     "Invalid (non-positive) sigma");
  double sig2 = sig * sig;
  double mode_u = mu - sig2;
  if (nsig == 0) {
    *out = exp(mode_u);
    return ORC_OK;
  }
  double isig = 1.0 / sig;
  double val_mode = prefac * isig * exp(0.5 * sig2 - mu);
  double target = val_mode * exp(-0.5 * nsig * nsig);
  orc2_lnp p = {mu, sig2, isig, isig * isig, target};
  double u_low = mode_u;
  double delta_u = log(exp(mode_u) + 2 * sig) - mode_u;
  double fval = orc2_lognorm(u_low + delta_u, &p);
  if (fval >= 0.0) {
    for (unsigned i = 0; i < max_bracket_iters; ++i) {
      delta_u *= 1.5;
      fval = orc2_lognorm(u_low + delta_u, &p);
      if (fval < 0.0) break;
    }
    if (fval >= 0.0)
      return orc2_fail(err, ORC_E_ARG, "numberCountsDoubleLogNormal", "logNormalSolver",
                       "Unable to bracket Log Normal");
  }
  double u_high = u_low + delta_u;
  orc2_brent s;
  double u = 0.0;
  if (orc2_brent_init(&s, &p, u_low, u_high, &u))
    return orc2_fail(err, ORC_E_ARG, "numberCountsDoubleLogNormal", "logNormalSolver",
                     "endpoints do not straddle y=0");
  int ok = 0;
  for (unsigned i = 0; i < max_bracket_iters; ++i) {
    orc2_brent_iterate(&s, &p, &u, &u_low, &u_high);
    if (orc2_test_interval(u_low, u_high, tol)) {
      ok = 1;
      break;
    }
  }
  if (!ok)
    return orc2_fail(err, ORC_E_ARG, "numberCountsDoubleLogNormal", "logNormalSolver",
                     "Unable to brent bracket Log Normal");
  *out = exp(u);
  return ORC_OK;
}

/* getMaxFlux (:1365-1380) */
int orc_ln_maxflux(const orc_lnmodel *m, double *out, orc_err *err) {
  const double nsig = 3.0;
  double kv = m->knots[m->nknots - 1];
  double sg = orc_ln_sigma(m, kv);
  double mu = orc_ln_offset(m, kv);
  double m2 = NAN;
  int rc = orc2_lognormal_solver(mu, sg, nsig, &m2, err);
  if (rc) return rc;
  out[0] = kv;
  out[1] = kv * m2;
  return ORC_OK;
}

/* getRRange / getRRangeInternal (:1481-1547, :1620-1632) */
int orc_ln_rrange(const orc_lnmodel *m, const orc_dbeam *bm, double *out,
                  orc_err *err) {
  if (!orc_ln_is_valid(m))
    return orc2_fail(err, ORC_E_INVALID_MODEL, "numberCountsDoubleLogNormal", "getRRange",
                     "Model is not valid");
  double mx[2];
  int rc = orc_ln_maxflux(m, mx, err);
  if (rc) return rc;
  double min1 = 0.0, max1 = 0.0, min2 = 0.0, max2 = 0.0, cval;
  if (bm->hassign[0]) {
    cval = mx[0] * bm->max1[0];
    if (cval > max1) max1 = cval;
    cval = mx[1] * bm->max2[0];
    if (cval > max2) max2 = cval;
  }
  if (bm->hassign[1]) {
    cval = mx[0] * bm->max1[1];
    if (cval > max1) max1 = cval;
    cval = -mx[1] * bm->max2[1];
    if (cval < min2) min2 = cval;
  }
  if (bm->hassign[2]) {
    cval = -mx[0] * bm->max1[2];
    if (cval < min1) min1 = cval;
    cval = mx[1] * bm->max2[2];
    if (cval > max2) max2 = cval;
  }
  if (bm->hassign[3]) {
    cval = -mx[0] * bm->max1[3];
    if (cval < min1) min1 = cval;
    cval = -mx[1] * bm->max2[3];
    if (cval < min2) min2 = cval;
  }
  out[0] = min1;
  out[1] = max1;
  out[2] = min2;
  out[3] = max2;
  return ORC_OK;
}

/* scalar getR (:1682-1727) */
double orc_ln_getR1(const orc_lnmodel *m, double f1, double f2,
                    const orc_dbeam *bm) {
  if (!orc_ln_is_valid(m)) return NAN;
  if ((f1 == 0) || (f2 == 0)) return 0.0;
  unsigned sgn = (f1 > 0) ? ((f2 > 0) ? 0 : 1) : ((f2 > 0) ? 2 : 3); /* signComp */
  if (!bm->hassign[sgn]) return 0.0;
  double af1 = fabs(f1), af2 = fabs(f2), retval = 0.0;
  if (bm->ishist[sgn]) {
    for (unsigned i = 0; i < bm->nhist[sgn]; ++i) {
      double ieta1 = bm->binv1[sgn][i], ieta2 = bm->binv2[sgn][i];
      retval += bm->binw[sgn][i] * ieta1 * ieta2 * orc2_counts_inner(m, af1 * ieta1, af2 * ieta2);
    }
  } else {
    for (unsigned i = 0; i < bm->npix[sgn]; ++i) {
      double ieta1 = bm->ipix1[sgn][i], ieta2 = bm->ipix2[sgn][i];
      retval += ieta1 * ieta2 * orc2_counts_inner(m, af1 * ieta1, af2 * ieta2);
    }
  }
  double prefac = bm->pixsize / 3600.0;
  return prefac * prefac * retval;
}

/* array getR (:1788-2007): R[i*n2+j]; includes (pixsize/3600)^2/sqrt(2pi) */
void orc_ln_getR(const orc_lnmodel *m, unsigned n1, const double *f1,
                 unsigned n2, const double *f2, const orc_dbeam *bm,
                 double *R) {
  if (!orc_ln_is_valid(m) || n1 == 0 || n2 == 0) {
    for (size_t i = 0; i < (size_t)n1 * n2; ++i) R[i] = NAN;
    return;
  }
  const double minknot = m->knots[0], maxknot = m->knots[m->nknots - 1];
  unsigned npsf[4];
  const double *wt[4], *ib1[4], *lr[4];
  double maxf1[4];
  for (int s = 0; s < 4; ++s) {
    npsf[s] = bm->hassign[s] ? (bm->ishist[s] ? bm->nhist[s] : bm->npix[s]) : 0;
    wt[s] = bm->ishist[s] ? bm->binw[s] : NULL;
    ib1[s] = bm->ishist[s] ? bm->binv1[s] : bm->ipix1[s];
    lr[s] = bm->ishist[s] ? bm->binlogratio[s] : bm->logratio[s];
    maxf1[s] = bm->hassign[s] ? maxknot * bm->max1[s] : 0.0;
  }
  unsigned maxnpsf = npsf[0] + npsf[1];
  if (npsf[2] + npsf[3] > maxnpsf) maxnpsf = npsf[2] + npsf[3];
  double *RW = (double *)malloc(sizeof(double) * 3 * (maxnpsf ? maxnpsf : 1));
  const double normfac = 1.0 / sqrt(2 * ORC2_PI);
  const double pixsizedeg = bm->pixsize / 3600.0;
  const double prefac = pixsizedeg * pixsizedeg * normfac;
  const double ilog2toe = 1.0 / log(2.0);
  const int hasNegX1 = bm->hassign[2] || bm->hassign[3];
  memset(R, 0, sizeof(double) * n1 * n2);
  unsigned *f2sgn = (unsigned *)malloc(sizeof(unsigned) * n2);
  double *f2inv = (double *)malloc(sizeof(double) * n2);
  double *f2log = (double *)malloc(sizeof(double) * n2);
  for (unsigned j = 0; j < n2; ++j) {
    double v = f2[j];
    if (v == 0) {
      f2inv[j] = 0.0;
      continue;
    }
    if (v > 0)
      f2sgn[j] = 0;
    else {
      f2sgn[j] = 1;
      v = fabs(v);
    }
    f2inv[j] = 1.0 / v;
    f2log[j] = log(v);
  }
  for (unsigned i = 0; i < n1; ++i) {
    double f1val = f1[i];
    if ((f1val == 0) || (f1val < 0 && !hasNegX1)) continue;
    unsigned sgn1;
    if (f1val > 0)
      sgn1 = 0;
    else {
      sgn1 = 2;
      f1val = fabs(f1val);
    }
    double logf1val = log(f1val);
    unsigned np = bm->hassign[sgn1] ? npsf[sgn1] : 0;
    unsigned nn = bm->hassign[sgn1 + 1] ? npsf[sgn1 + 1] : 0;
    for (unsigned k = 0; k < 2; ++k) {
      unsigned sgn = sgn1 + k;
      unsigned curr_n = (k == 0) ? np : nn;
      double *RWp = RW + k * np * 3;
      if (curr_n > 0 && f1val < maxf1[sgn]) {
        for (unsigned j = 0; j < curr_n; ++j) {
          double ieta1 = ib1[sgn][j];
          double f1prod = f1val * ieta1;
          double *c = RWp + 3 * j;
          if (f1prod >= minknot && f1prod < maxknot) {
            double cts = exp2(orc_cspline_eval(m->nknots, m->logknots, m->logknotvals,
                                               m->kc, log2(f1prod)));
            if (cts > 0) {
              double isigma = 1.0 / orc_ln_sigma(m, f1prod);
              if (bm->ishist[sgn])
                c[0] = wt[sgn][j] * ieta1 * isigma * cts * prefac;
              else
                c[0] = cts * ieta1 * isigma * prefac;
              c[1] = logf1val + orc_ln_offset(m, f1prod) - lr[sgn][j];
              c[2] = -0.5 * isigma * isigma * ilog2toe;
            } else
              c[0] = 0;
          } else
            c[0] = 0;
        }
      } else {
        unsigned lo = (k == 0) ? 0 : np, hi = (k == 0) ? np : (np + nn);
        for (unsigned j = lo; j < hi; ++j) RW[3 * j] = 0.0;
      }
    }
    double *row = R + (size_t)i * n2;
    for (unsigned j = 0; j < n2; ++j) {
      double if2 = f2inv[j];
      if (if2 == 0) continue;
      unsigned fs = f2sgn[j];
      unsigned sgn = sgn1 + fs;
      if (bm->hassign[sgn]) {
        double workval = 0;
        unsigned curr_n = npsf[sgn];
        const double *RWp = RW + fs * np * 3;
        double lf2 = f2log[j];
        for (unsigned k = 0; k < curr_n; ++k) {
          const double *c = RWp + k * 3;
          double c0 = c[0];
          if (c0 != 0) {
            double tfac = lf2 - c[1];
            workval += c0 * exp2(tfac * tfac * c[2]);
          }
        }
        row[j] = workval * f2inv[j];
      }
    }
  }
  free(RW);
  free(f2sgn);
  free(f2inv);
  free(f2log);
}

/* ===================================================================== */
/* fitsDataDouble (src/fitsDataDouble.cc)                                */
/* ===================================================================== */

/* getMinMax (:481-512, unbinned) */
void orc_data2_minmax(const double *d1, const double *d2, size_t n, double *out) {
  double mn1 = d1[0], mx1 = d1[0], mn2 = d2[0], mx2 = d2[0];
  for (size_t i = 1; i < n; ++i) {
    if (d1[i] < mn1) mn1 = d1[i];
    if (d1[i] > mx1) mx1 = d1[i];
  }
  for (size_t i = 1; i < n; ++i) {
    if (d2[i] < mn2) mn2 = d2[i];
    if (d2[i] > mx2) mx2 = d2[i];
  }
  out[0] = mn1;
  out[1] = mx1;
  out[2] = mn2;
  out[3] = mx2;
}

/* getMean + meanSubtract (:517-552) */
void orc_data2_mean_subtract(double *d1, double *d2, size_t n, double *mean) {
  double m = d1[0];
  for (size_t i = 1; i < n; ++i) m += d1[i];
  mean[0] = m / (double)n;
  m = d2[0];
  for (size_t i = 1; i < n; ++i) m += d2[i];
  mean[1] = m / (double)n;
  for (size_t i = 0; i < n; ++i) d1[i] -= mean[0];
  for (size_t i = 0; i < n; ++i) d2[i] -= mean[1];
}

/* applyBinning (:557-608) */
int orc_data2_bin(const double *d1, const double *d2, size_t n, uint32_t nb1,
                  uint32_t nb2, uint32_t *bins, double *out) {
  if (nb1 == 0 || nb2 == 0) return ORC_E_ARG;
  memset(bins, 0, sizeof(uint32_t) * nb1 * nb2);
  double mm[4];
  orc_data2_minmax(d1, d2, n, mm);
  for (int i = 0; i < 4; ++i)
    if (isnan(mm[i]) || isinf(mm[i])) return ORC_E_ARG;
  double c01 = mm[0], c02 = mm[2], bd1, bd2;
  bd1 = (nb1 == 1) ? 2 * (mm[1] - mm[0]) : (mm[1] - mm[0]) / (double)(nb1 - 1);
  bd2 = (nb2 == 1) ? 2 * (mm[3] - mm[2]) : (mm[3] - mm[2]) / (double)(nb2 - 1);
  double ib1 = 1.0 / bd1, ib2 = 1.0 / bd2;
  for (size_t i = 0; i < n; ++i) {
    unsigned idx1 = (unsigned)((d1[i] - c01) * ib1 + 0.5);
    unsigned idx2 = (unsigned)((d2[i] - c02) * ib2 + 0.5);
    bins[(size_t)idx1 * nb2 + idx2] += 1;
  }
  out[0] = c01;
  out[1] = bd1;
  out[2] = c02;
  out[3] = bd2;
  return ORC_OK;
}

/* ===================================================================== */
/* PDDouble (src/PDDouble.cc)                                            */
/* ===================================================================== */

/* getIntegral (:142-188) */
double orc_pd2_integral(unsigned n1, unsigned n2, const double *pd,
                        double dflux1, double dflux2, int islog) {
#define V(x) (islog ? exp2(x) : (x))
  double tot = 0.5 * V(pd[0]);
  for (unsigned j = 1; j < n2 - 1; ++j) tot += V(pd[j]);
  tot += 0.5 * V(pd[n2 - 1]);
  tot *= 0.5;
  for (unsigned i = 1; i < n1 - 1; ++i) {
    const double *r = pd + (size_t)i * n2;
    tot += 0.5 * V(r[0]);
    for (unsigned j = 1; j < n2 - 1; ++j) tot += V(r[j]);
    tot += 0.5 * V(r[n2 - 1]);
  }
  const double *r = pd + (size_t)(n1 - 1) * n2;
  tot += 0.25 * V(r[0]);
  for (unsigned j = 1; j < n2 - 1; ++j) tot += 0.5 * V(r[j]);
  tot += 0.25 * V(r[n2 - 1]);
#undef V
  return tot * dflux1 * dflux2;
}

/* getMeans (:355-452) */
void orc_pd2_means(unsigned n1, unsigned n2, const double *pd, double minflux1,
                   double dflux1, double minflux2, double dflux2, int islog,
                   int donorm, double *out) {
#define V(x) (islog ? exp2(x) : (x))
  double mean1 = 0.0, mean2 = 0.0, xsum, pval;
  for (unsigned j = 1; j < n2 - 1; ++j) mean2 += j * V(pd[j]);
  mean2 += 0.5 * (n2 - 1) * V(pd[n2 - 1]);
  mean2 *= 0.5;
  for (unsigned i = 1; i < n1 - 1; ++i) {
    const double *r = pd + (size_t)n2 * i;
    xsum = 0.5 * V(r[0]);
    for (unsigned j = 1; j < n2 - 1; ++j) {
      pval = V(r[j]);
      xsum += pval;
      mean2 += (double)j * pval;
    }
    pval = V(r[n2 - 1]);
    xsum += 0.5 * pval;
    mean1 += xsum * (double)i;
    mean2 += 0.5 * (n2 - 1) * pval;
  }
  const double *r = pd + (size_t)(n1 - 1) * n2;
  xsum = 0.5 * V(r[0]);
  for (unsigned j = 1; j < n2 - 1; ++j) {
    pval = V(r[j]);
    xsum += pval;
    mean2 += 0.5 * (double)j * pval;
  }
  pval = V(r[n2 - 1]);
  xsum += 0.5 * pval;
  mean1 += 0.5 * xsum * (n1 - 1);
  mean2 += 0.25 * (n2 - 1) * pval;
#undef V
  mean1 *= dflux1 * dflux1 * dflux2;
  mean2 *= dflux1 * dflux2 * dflux2;
  if (donorm) {
    double inorm = 1.0 / orc_pd2_integral(n1, n2, pd, dflux1, dflux2, islog);
    mean1 *= inorm;
    mean2 *= inorm;
  }
  out[0] = mean1 + minflux1;
  out[1] = mean2 + minflux2;
}

/* a P(D) value in log2: as stored (logflat) or log2 of it */
static double orc2_lv(const double *pd, size_t i, int islog) {
  return islog ? pd[i] : log2(pd[i]);
}

/* PDDouble::getLogLikeUnbinned (src/PDDouble.cc:994-1082): the log2-stored
   branch tests the edges with <= / >=, the linear-stored one with < / >
   (:1047-1066) and takes log2 of each value it reads */
double orc_pd2_loglike_unbinned(unsigned n1, unsigned n2, const double *pd,
                                double minflux1, double dflux1,
                                double minflux2, double dflux2, int islog,
                                const double *d1, const double *d2,
                                size_t ndata) {
  double maxflux1 = minflux1 + (double)(n1 - 1) * dflux1;
  double maxflux2 = minflux2 + (double)(n2 - 1) * dflux2;
  unsigned n2n1 = n2 * n1;
  double idflux1 = 1.0 / dflux1, idflux2 = 1.0 / dflux2, loglike = 0.0;
  for (size_t i = 0; i < ndata; ++i) {
    double cflux1 = d1[i], cflux2 = d2[i], iv;
    double delt1 = (cflux1 - minflux1) * idflux1;
    double delt2 = (cflux2 - minflux2) * idflux2;
    int idx1 = (int)delt1, idx2 = (int)delt2;
    int n2idx1 = n2 * idx1;
    int lo1 = islog ? (cflux1 <= minflux1) : (cflux1 < minflux1);
    int hi1 = islog ? (cflux1 >= maxflux1) : (cflux1 > maxflux1);
    int lo2 = islog ? (cflux2 <= minflux2) : (cflux2 < minflux2);
    int hi2 = islog ? (cflux2 >= maxflux2) : (cflux2 > maxflux2);
    if (lo1) {
      if (lo2) iv = orc2_lv(pd, 0, islog);
      else if (hi2) iv = orc2_lv(pd, n2 - 1, islog);
      else iv = orc2_lv(pd, idx2, islog);
    } else if (hi1) {
      if (lo2) iv = orc2_lv(pd, n2n1 - n1, islog);
      else if (hi2) iv = orc2_lv(pd, n2n1 - 1, islog);
      else iv = orc2_lv(pd, n2n1 - n1 + idx2, islog);
    } else if (lo2) {
      iv = orc2_lv(pd, n2idx1, islog);
    } else if (hi2) {
      iv = orc2_lv(pd, n2idx1 + n2 - 1, islog);
    } else {
      double t = delt1 - (double)idx1, u = delt2 - (double)idx2;
      double omu = 1.0 - u, omt = 1.0 - t;
      unsigned base = n2idx1 + idx2;
      iv = omt * (omu * orc2_lv(pd, base, islog) + u * orc2_lv(pd, base + 1, islog)) +
           t * (omu * orc2_lv(pd, base + n2, islog) + u * orc2_lv(pd, base + n2 + 1, islog));
    }
    loglike += iv;
  }
  return log(2.) * loglike;
}

/* PDDouble::getLogLikeBinned (src/PDDouble.cc:1092-1220): both branches
   test the edges with < / >; the linear-stored one takes log2 of each
   value it reads */
double orc_pd2_loglike_binned(unsigned n1, unsigned n2, const double *pd,
                              double minflux1, double dflux1, double minflux2,
                              double dflux2, int islog, const uint32_t *bins,
                              uint32_t nb1, uint32_t nb2, double bincent01,
                              double bindelta1, double bincent02,
                              double bindelta2) {
  double maxflux1 = minflux1 + (double)(n1 - 1) * dflux1;
  double maxflux2 = minflux2 + (double)(n2 - 1) * dflux2;
  unsigned n2n1 = n2 * n1;
  double idflux1 = 1.0 / dflux1, idflux2 = 1.0 / dflux2, loglike = 0.0;
  for (unsigned i = 0; i < nb1; ++i) {
    double cflux1 = bincent01 + (double)i * bindelta1;
    int idx1 = (int)((cflux1 - minflux1) * idflux1);
    int n2idx1 = n2 * idx1;
    const uint32_t *bp = bins + (size_t)i * nb2;
    for (unsigned j = 0; j < nb2; ++j) {
      uint32_t ninbin = bp[j];
      if (ninbin == 0) continue;
      double cflux2 = bincent02 + (double)j * bindelta2, iv;
      int idx2 = (int)((cflux2 - minflux2) * idflux2);
      if (cflux1 < minflux1) {
        if (cflux2 < minflux2) iv = orc2_lv(pd, 0, islog);
        else if (cflux2 > maxflux2) iv = orc2_lv(pd, n2 - 1, islog);
        else iv = orc2_lv(pd, idx2, islog);
      } else if (cflux1 > maxflux1) {
        if (cflux2 < minflux2) iv = orc2_lv(pd, n2n1 - n1, islog);
        else if (cflux2 > maxflux2) iv = orc2_lv(pd, n2n1 - 1, islog);
        else iv = orc2_lv(pd, n2n1 - n1 + idx2, islog);
      } else if (cflux2 < minflux2) {
        iv = orc2_lv(pd, n2idx1, islog);
      } else if (cflux2 > maxflux2) {
        iv = orc2_lv(pd, n2idx1 + n2 - 1, islog);
      } else {
        double t = (cflux1 - minflux1) * idflux1 - (double)idx1;
        double u = (cflux2 - minflux2) * idflux2 - (double)idx2;
        double omu = 1.0 - u, omt = 1.0 - t;
        unsigned base = n2idx1 + idx2;
        iv = omt * (omu * orc2_lv(pd, base, islog) + u * orc2_lv(pd, base + 1, islog)) +
             t * (omu * orc2_lv(pd, base + n2, islog) + u * orc2_lv(pd, base + n2 + 1, islog));
      }
      loglike += (double)ninbin * iv;
    }
  }
  return log(2.) * loglike;
}

/* ===================================================================== */
/* PDFactoryDouble (src/PDFactoryDouble.cc)                              */
/* ===================================================================== */

typedef struct orc_pdfac2 {
  unsigned n, nedge;
  double *RFlux1, *RFlux2;
  unsigned wrap1, wrap2;
  double dflux1, dflux2;
  double *rvals, *prtrans, *pval, *pofd, *rsum, *pd;
  double *edge1, *edge2, *rxw, *ryw, *edgework;
  double mn1, mn2, var1, var2;
  double sg1pos, sg1neg, sg2pos, sg2neg;
  int doshift1, doshift2;
  double shift1, shift2, r0;
  unsigned split1, split2;
  double pd_minflux1, pd_minflux2;
} orc_pdfac2;

static orc_pdfac2 *orc_pdfac2_new(unsigned n, unsigned nedge) {
  orc_pdfac2 *p = (orc_pdfac2 *)calloc(1, sizeof(orc_pdfac2));
  p->n = n;
  p->nedge = nedge;
  const size_t nn = (size_t)n * n, nc = (size_t)n * (n / 2 + 1);
  p->RFlux1 = (double *)malloc(sizeof(double) * n);
  p->RFlux2 = (double *)malloc(sizeof(double) * n);
  p->rvals = (double *)malloc(sizeof(double) * nn);
  p->prtrans = (double *)malloc(sizeof(double) * 2 * nc);
  p->pval = (double *)malloc(sizeof(double) * 2 * nc);
  p->pofd = (double *)malloc(sizeof(double) * nn);
  p->pd = (double *)malloc(sizeof(double) * nn);
  p->rsum = (double *)malloc(sizeof(double) * n);
  p->edge1 = (double *)malloc(sizeof(double) * nedge);
  p->edge2 = (double *)malloc(sizeof(double) * nedge);
  p->rxw = (double *)malloc(sizeof(double) * (n > nedge ? n : nedge));
  p->ryw = (double *)malloc(sizeof(double) * (n > nedge ? n : nedge));
  size_t ew = (size_t)nedge * (n > nedge ? n : nedge);
  p->edgework = (double *)malloc(sizeof(double) * ew);
  return p;
}

static void orc_pdfac2_free(orc_pdfac2 *p) {
  if (!p) return;
  free(p->RFlux1);
  free(p->RFlux2);
  free(p->rvals);
  free(p->prtrans);
  free(p->pval);
  free(p->pofd);
  free(p->pd);
  free(p->rsum);
  free(p->edge1);
  free(p->edge2);
  free(p->rxw);
  free(p->ryw);
  free(p->edgework);
  free(p);
}

/* initRFluxInternal (:18-48) */
static void orc2_init_rflux(unsigned n, double minflux, double maxflux,
                            double *rflux, double *dflux, unsigned *wrapidx) {
  double inm1 = 1.0 / (double)(n - 1);
  double df = (maxflux - minflux) * inm1;
  if (minflux >= 0.0) {
    rflux[0] = minflux;
    for (unsigned i = 1; i < n; ++i) rflux[i] = (double)i * df + minflux;
    *wrapidx = n - 1;
  } else {
    df = maxflux / (n - floor(-minflux / df) - 1.0);
    unsigned w = (unsigned)(maxflux / df + 0.999999999);
    rflux[0] = 0.0;
    for (unsigned i = 1; i <= w; ++i) rflux[i] = (double)i * df;
    double wrapval = -(double)n * df;
    for (unsigned i = w + 1; i < n; ++i) rflux[i] = (double)i * df + wrapval;
    *wrapidx = w;
  }
  *dflux = df;
}

/* initR with edges (:355-551), muldflux = true */
static void orc2_init_r(orc_pdfac2 *p, double minflux1, double maxflux1,
                        double minflux2, double maxflux2,
                        const orc_lnmodel *model, const orc_dbeam *bm,
                        int setEdge) {
  const unsigned n = p->n, nedge = p->nedge;
  if (maxflux1 < minflux1) {
    double t = minflux1;
    minflux1 = maxflux1;
    maxflux1 = t;
  }
  if (maxflux2 < minflux2) {
    double t = minflux2;
    minflux2 = maxflux2;
    maxflux2 = t;
  }
  orc2_init_rflux(n, minflux1, maxflux1, p->RFlux1, &p->dflux1, &p->wrap1);
  orc2_init_rflux(n, minflux2, maxflux2, p->RFlux2, &p->dflux2, &p->wrap2);
  const double dflux1 = p->dflux1, dflux2 = p->dflux2;
  double r00val = 0.0;
  if (setEdge) {
    const double lowEdgeRMult = 1e-9;
    double minedge1 = dflux1 * lowEdgeRMult;
    double minedge2 = dflux2 * lowEdgeRMult;
    double inedgem1 = 1.0 / (double)(nedge - 1);
    /* use_edge_log_x = true, use_edge_log_y = false (PDFactoryDouble.h:33-37) */
    double dife1 = -log(lowEdgeRMult) * inedgem1;
    for (unsigned i = 0; i < nedge; ++i) p->edge1[i] = minedge1 * exp((double)i * dife1);
    double dife2 = (dflux2 - minedge2) * inedgem1;
    for (unsigned i = 0; i < nedge; ++i) p->edge2[i] = minedge2 + (double)i * dife2;
    double iRxnorm = dife1 / (dflux1 - minedge1);
    double iRynorm = dife2 / (dflux2 - minedge2);
    double iR00norm = dife1 * dife2 / ((dflux1 - minedge1) * (dflux2 - minedge2));
    double *rv = p->edgework, s;
    /* R[0,0]: y integral (linear) then x integral (log) */
    orc_ln_getR(model, nedge, p->edge1, nedge, p->edge2, bm, rv);
    for (unsigned i = 0; i < nedge; ++i) {
      const double *r = rv + (size_t)i * nedge;
      s = 0.5 * r[0];
      for (unsigned j = 1; j < nedge - 1; ++j) s += r[j];
      s += 0.5 * r[nedge - 1];
      p->rxw[i] = s;
    }
    s = 0.5 * p->edge1[0] * p->rxw[0];
    for (unsigned i = 1; i < nedge - 1; ++i) s += p->edge1[i] * p->rxw[i];
    s += 0.5 * p->edge1[nedge - 1] * p->rxw[nedge - 1];
    r00val = s * iR00norm;
    /* Rx = R[0, y]: x integral (log) at every RFlux2 */
    orc_ln_getR(model, nedge, p->edge1, n, p->RFlux2, bm, rv);
    double rf = p->edge1[0];
    for (unsigned j = 0; j < n; ++j) p->rxw[j] = 0.5 * rf * rv[j];
    for (unsigned i = 1; i < nedge - 1; ++i) {
      rf = p->edge1[i];
      const double *r = rv + (size_t)i * n;
      for (unsigned j = 0; j < n; ++j) p->rxw[j] += rf * r[j];
    }
    rf = p->edge1[nedge - 1];
    {
      const double *r = rv + (size_t)(nedge - 1) * n;
      for (unsigned j = 0; j < n; ++j) p->rxw[j] += 0.5 * rf * r[j];
    }
    for (unsigned j = 0; j < n; ++j) p->rxw[j] *= iRxnorm;
    /* Ry = R[x, 0]: y integral (linear) at every RFlux1 */
    orc_ln_getR(model, n, p->RFlux1, nedge, p->edge2, bm, rv);
    for (unsigned i = 0; i < n; ++i) {
      const double *r = rv + (size_t)i * nedge;
      s = 0.5 * r[0];
      for (unsigned j = 1; j < nedge - 1; ++j) s += r[j];
      s += 0.5 * r[nedge - 1];
      p->ryw[i] = s;
    }
    for (unsigned i = 0; i < n; ++i) p->ryw[i] *= iRynorm;
  }
  orc_ln_getR(model, n, p->RFlux1, n, p->RFlux2, bm, p->rvals);
  if (setEdge) {
    memcpy(p->rvals + 1, p->rxw + 1, (n - 1) * sizeof(double));
    for (unsigned i = 1; i < n; ++i) p->rvals[(size_t)i * n] = p->ryw[i];
    p->rvals[0] = r00val;
  } else {
    memset(p->rvals, 0, n * sizeof(double));
    for (unsigned i = 1; i < n; ++i) p->rvals[(size_t)i * n] = 0.0;
  }
  const double fluxfac = dflux1 * dflux2;
  for (size_t i = 0; i < (size_t)n * n; ++i) p->rvals[i] *= fluxfac;
}

/* getRStats (:593-722), rdflux = true */
static void orc2_rstats(orc_pdfac2 *p) {
  const unsigned cs = p->n;
  const double *RF1 = p->RFlux1, *RF2 = p->RFlux2;
  double cf1, cf2, cR, mean1, mean2, var1, var2, prod1, prod2;
  const double *row = p->rvals;
  cf1 = RF1[0];
  cR = 0.5 * row[0];
  prod1 = 0.5 * cf1 * cR;
  mean1 = prod1;
  var1 = cf1 * prod1;
  cf2 = RF2[0];
  prod2 = 0.5 * cf2 * cR;
  mean2 = prod2;
  var2 = cf2 * prod2;
  for (unsigned j = 1; j < cs - 1; ++j) {
    cR = 0.5 * row[j];
    prod1 = cf1 * cR;
    mean1 += prod1;
    var1 += cf1 * prod1;
    cf2 = RF2[j];
    prod2 = cf2 * cR;
    mean2 += prod2;
    var2 += cf2 * prod2;
  }
  cR = 0.5 * row[cs - 1];
  prod1 = 0.5 * cf1 * cR;
  mean1 += prod1;
  var1 += cf1 * prod1;
  cf2 = RF2[cs - 1];
  prod2 = 0.5 * cf2 * cR;
  mean2 += prod2;
  var2 += cf2 * prod2;
  for (unsigned i = 1; i < cs - 1; ++i) {
    row = p->rvals + (size_t)i * cs;
    cf1 = RF1[i];
    cR = row[0];
    prod1 = 0.5 * cf1 * cR;
    mean1 += prod1;
    var1 += cf1 * prod1;
    cf2 = RF2[0];
    prod2 = 0.5 * cf2 * cR;
    mean2 += prod2;
    var2 += cf2 * prod2;
    for (unsigned j = 1; j < cs - 1; ++j) {
      cR = row[j];
      prod1 = cf1 * cR;
      mean1 += prod1;
      var1 += cf1 * prod1;
      cf2 = RF2[j];
      prod2 = cf2 * cR;
      mean2 += prod2;
      var2 += cf2 * prod2;
    }
    cR = row[cs - 1];
    prod1 = 0.5 * cf1 * cR;
    mean1 += prod1;
    var1 += cf1 * prod1;
    cf2 = RF2[cs - 1];
    prod2 = 0.5 * cf2 * cR;
    mean2 += prod2;
    var2 += cf2 * prod2;
  }
  row = p->rvals + (size_t)(cs - 1) * cs;
  cf1 = RF1[cs - 1];
  cR = 0.5 * row[0];
  prod1 = 0.5 * cf1 * cR;
  mean1 += prod1;
  var1 += cf1 * prod1;
  cf2 = RF2[0];
  prod2 = 0.5 * cf2 * cR;
  mean2 += prod2;
  var2 += cf2 * prod2;
  for (unsigned j = 1; j < cs - 1; ++j) {
    cR = 0.5 * row[j];
    prod1 = cf1 * cR;
    mean1 += prod1;
    var1 += cf1 * prod1;
    cf2 = RF2[j];
    prod2 = cf2 * cR;
    mean2 += prod2;
    var2 += cf2 * prod2;
  }
  cR = 0.5 * row[cs - 1];
  prod1 = 0.5 * cf1 * cR;
  mean1 += prod1;
  var1 += cf1 * prod1;
  cf2 = RF2[cs - 1];
  prod2 = 0.5 * cf2 * cR;
  mean2 += prod2;
  var2 += cf2 * prod2;
  p->mn1 = mean1;
  p->mn2 = mean2;
  p->var1 = var1;
  p->var2 = var2;
}

/* initPD (:1145-1342) */
static int orc2_initpd(orc_pdfac2 *p, double minflux1, double maxflux1,
                       double minflux2, double maxflux2,
                       const orc_lnmodel *model, const orc_dbeam *bm,
                       int setEdge, orc_err *err) {
  const unsigned n = p->n;
  if (maxflux1 <= 0.0 || maxflux2 <= 0.0)
    return orc2_fail(err, ORC_E_ARG, "PDFactoryDouble", "initPD", "Invalid (non-positive) maxflux");
  if (maxflux1 == minflux1 || maxflux2 == minflux2)
    return orc2_fail(err, ORC_E_ARG, "PDFactoryDouble", "initPD", "Invalid range (0) between min/max flux");
  orc2_init_r(p, minflux1, maxflux1, minflux2, maxflux2, model, bm, setEdge);
  orc2_rstats(p);
  p->sg1pos = sqrt(p->var1);
  p->sg2pos = sqrt(p->var2);
  p->doshift1 = (p->sg1pos > p->dflux1) && (fabs(p->mn1) > p->dflux1);
  p->shift1 = p->doshift1 ? -p->mn1 : 0.0;
  p->doshift2 = (p->sg2pos > p->dflux2) && (fabs(p->mn2) > p->dflux2);
  p->shift2 = p->doshift2 ? -p->mn2 : 0.0;
  orc_rfft2(n, p->rvals, p->prtrans);
  const unsigned nc = n / 2 + 1;
  const double iflux1 = orc2_two_pi / (n * p->dflux1);
  const double iflux2 = orc2_two_pi / (n * p->dflux2);
  double r0 = p->prtrans[0];
  p->r0 = r0;
  const double sf1 = p->doshift1 ? p->shift1 * iflux1 : 0.0;
  const double sf2 = p->doshift2 ? p->shift2 * iflux2 : 0.0;
  for (unsigned i1 = 0; i1 < n; ++i1) {
    double *row = p->prtrans + 2 * (size_t)i1 * nc;
    double sp1 = 0.0;
    if (p->doshift1) sp1 = (i1 < nc) ? sf1 * (double)i1 : -sf1 * (double)(n - i1);
    for (unsigned i2 = 0; i2 < nc; ++i2) {
      double rval = row[2 * i2] - r0;
      double ival = row[2 * i2 + 1];
      if (p->doshift1 && p->doshift2)
        ival = ival - sp1 - sf2 * (double)i2;
      else if (p->doshift1)
        ival = ival - sp1;
      else if (p->doshift2)
        ival = ival - sf2 * (double)i2;
      double expfac = exp(rval);
      row[2 * i2] = expfac * cos(ival);
      row[2 * i2 + 1] = expfac * sin(ival);
    }
  }
  p->prtrans[0] = 1.0;
  p->prtrans[1] = 0.0;
  return ORC_OK;
}

/* findSplitPoint (:784-975) */
static int orc2_find_split(const double *data, unsigned cs, double dflux,
                           double sigmapos, double sigmaneg, unsigned *out,
                           orc_err *err) {
  const double n1 = 2.25, f1 = 5e-14, f2 = 1.5e-5, f3 = 5;
  if (sigmapos <= 0 || sigmaneg <= 0)
    return orc2_fail(err, ORC_E_SPLIT, "PDFactory", "findSplitPoint", "Invalid (non-positive) sigma");
  double fluxrange = cs * dflux;
  double n1sigpos = n1 * sigmapos, n1signeg = n1 * sigmaneg;
  if (fluxrange < (n1sigpos + n1signeg))
    return orc2_fail(err, ORC_E_SPLIT, "PDFactoryDouble", "findSplitPoint", "Not enough flux range");
  double peak = data[0];
  for (unsigned i = 1; i < cs; ++i)
    if (data[i] > peak) peak = data[i];
  if (peak <= 0.0)
    return orc2_fail(err, ORC_E_SPLIT, "PDFactoryDouble", "findSplitPoint", "Peak P(D) value was zero");
  int topidx = (int)cs - (int)(n1signeg / dflux) - 1;
  int botidx = (int)(n1sigpos / dflux);
  if (topidx >= (int)cs || botidx < 0 || botidx > topidx)
    return orc2_fail(err, ORC_E_SPLIT, "PDFactoryDouble", "findSplitPoint", "Logic error in topidx/botidx");
  int minidx = botidx;
  double minval = data[minidx];
  for (int i = botidx + 1; i <= topidx; ++i)
    if (data[i] < minval) {
      minidx = i;
      minval = data[i];
    }
  if (minval > f2 * peak)
    return orc2_fail(err, ORC_E_SPLIT, "PDFactoryDouble", "findSplitPoint",
                     "Minimum in marginalized pd is too large");
  const int pre_jitter_n = 17;
  double jitter = 0.0;
  if (minidx > pre_jitter_n) {
    int mono = 1;
    int initidx = minidx - pre_jitter_n;
    double prev = data[initidx];
    jitter = prev;
    for (int i = initidx + 1; i < minidx; ++i) {
      double cv = data[i];
      if (cv >= prev) mono = 0;
      jitter += cv;
      prev = cv;
    }
    if (mono) {
      *out = (unsigned)minidx;
      return ORC_OK;
    }
    jitter += minval;
    jitter = 1.1 * jitter / (double)(pre_jitter_n + 1);
  } else {
    jitter = f1 * peak;
    if (minval > jitter) {
      *out = (unsigned)minidx;
      return ORC_OK;
    }
  }
  if (minval <= 0.0) {
    const int nminup = 3;
    int bi, ti;
    if (minidx < botidx)
      bi = botidx;
    else
      bi = (minidx >= nminup) ? (minidx - nminup) : minidx;
    if (minidx > topidx - nminup)
      ti = topidx - nminup;
    else
      ti = minidx + nminup;
    ti = (ti >= (int)cs) ? ((int)cs - 1) : ti;
    for (int i = bi; i < ti; ++i)
      if (data[i] > minval) minval = data[i];
  }
  if (minval > 0) jitter = (f3 * minval) < jitter ? (f3 * minval) : jitter;
  if ((botidx >= topidx) || (abs(topidx - botidx) < 1))
    return orc2_fail(err, ORC_E_SPLIT, "PDFactoryDouble", "findSplitPoint", "Topidx vs. botidx search issue");
  if (data[topidx] < jitter)
    return orc2_fail(err, ORC_E_SPLIT, "PDFactoryDouble", "findSplitPoint", "Topidx is already below jitter");
  for (int i = topidx - 1; i >= botidx; --i)
    if (data[i] < jitter) {
      *out = (unsigned)(i + 1);
      return ORC_OK;
    }
  return orc2_fail(err, ORC_E_SPLIT, "PDFactoryDouble", "findSplitPoint", "Unable to find split point");
}

/* findSplitPoint on a caller's marginal (hand-built cases) */
int orc2_find_split_array(const double *data, unsigned cs, double dflux,
                          double sigmapos, double sigmaneg, unsigned *out,
                          orc_err *err) {
  return orc2_find_split(data, cs, dflux, sigmapos, sigmaneg, out, err);
}

/* unwrapAndNormalizePD (:978-1118) into p->pd */
static int orc2_unwrap_normalize(orc_pdfac2 *p, orc_err *err) {
  const unsigned cs = p->n;
  double *pofd = p->pofd;
  for (size_t i = 0; i < (size_t)cs * cs; ++i)
    if (pofd[i] < 0) pofd[i] = 0.0;
  for (unsigned i = 0; i < cs; ++i) {
    const double *r = pofd + (size_t)i * cs;
    double c = 0.5 * r[0];
    for (unsigned j = 1; j < cs - 1; ++j) c += r[j];
    c += 0.5 * r[cs - 1];
    p->rsum[i] = c;
  }
  unsigned s1, s2;
  int rc = orc2_find_split(p->rsum, cs, p->dflux1, p->sg1pos, p->sg1neg, &s1, err);
  if (rc) return rc;
  for (unsigned j = 0; j < cs; ++j) p->rsum[j] = 0.5 * pofd[j];
  for (unsigned i = 1; i < cs - 1; ++i) {
    const double *r = pofd + (size_t)i * cs;
    for (unsigned j = 0; j < cs; ++j) p->rsum[j] += r[j];
  }
  {
    const double *r = pofd + (size_t)(cs - 1) * cs;
    for (unsigned j = 0; j < cs; ++j) p->rsum[j] += 0.5 * r[j];
  }
  rc = orc2_find_split(p->rsum, cs, p->dflux2, p->sg2pos, p->sg2neg, &s2, err);
  if (rc) return rc;
  p->split1 = s1;
  p->split2 = s2;
  /* the four-quadrant copy is the cyclic shift by (s1, s2) */
  for (unsigned i = 0; i < cs; ++i) {
    unsigned si = (i + s1) % cs;
    for (unsigned j = 0; j < cs; ++j) p->pd[(size_t)i * cs + j] = pofd[(size_t)si * cs + (j + s2) % cs];
  }
  double tot = orc_pd2_integral(cs, cs, p->pd, p->dflux1, p->dflux2, 0);
  double itot = 1.0 / tot;
  for (size_t i = 0; i < (size_t)cs * cs; ++i) p->pd[i] *= itot;
  double mn[2];
  orc_pd2_means(cs, cs, p->pd, 0.0, p->dflux1, 0.0, p->dflux2, 0, 0, mn);
  if (isinf(mn[0]) || isnan(mn[0]) || isinf(mn[1]) || isnan(mn[1]))
    return orc2_fail(err, ORC_E_MEAN, "PDFactoryDouble", "unwrapAndNormalizePD",
                     "Un-shift amounts not finite");
  p->pd_minflux1 = -mn[0];
  p->pd_minflux2 = -mn[1];
  return ORC_OK;
}

/* getPD (:1362-1460) */
static int orc2_getpd(orc_pdfac2 *p, double sigma1, double sigma2, int setlog,
                      orc_err *err) {
  const unsigned n = p->n, nc = n / 2 + 1;
  if (sigma1 > 0 || sigma2 > 0) {
    double iflux1 = orc2_two_pi / (n * p->dflux1);
    double iflux2 = orc2_two_pi / (n * p->dflux2);
    double sigfac1 = -0.5 * sigma1 * sigma1 * iflux1 * iflux1;
    double sigfac2 = -0.5 * sigma2 * sigma2 * iflux2 * iflux2;
    for (unsigned i1 = 0; i1 < n; ++i1) {
      double didx1 = (i1 < nc) ? (double)i1 : (double)(n - i1);
      double sigprod1 = sigfac1 * didx1 * didx1;
      const double *in = p->prtrans + 2 * (size_t)i1 * nc;
      double *out = p->pval + 2 * (size_t)i1 * nc;
      for (unsigned i2 = 0; i2 < nc; ++i2) {
        double didx2 = (double)i2;
        double expfac = exp(sigprod1 + sigfac2 * didx2 * didx2);
        out[2 * i2] = expfac * in[2 * i2];
        out[2 * i2 + 1] = expfac * in[2 * i2 + 1];
      }
    }
    p->pval[0] = 1.0;
    p->pval[1] = 0.0;
    orc_irfft2(n, p->pval, p->pofd);
  } else {
    orc_irfft2(n, p->prtrans, p->pofd);
  }
  p->sg1pos = sqrt(p->var1 + sigma1 * sigma1);
  p->sg2pos = sqrt(p->var2 + sigma2 * sigma2);
  p->sg1neg = (9 * sigma1 * sigma1 > p->var1) ? sigma1 : p->sg1pos;
  p->sg2neg = (9 * sigma2 * sigma2 > p->var2) ? sigma2 : p->sg2pos;
  int rc = orc2_unwrap_normalize(p, err);
  if (rc) return rc;
  if (setlog) {
    for (size_t i = 0; i < (size_t)n * n; ++i) {
      double v = p->pd[i];
      p->pd[i] = (v <= 0.0) ? orc2_smalllogval : log2(v);
    }
  }
  return ORC_OK;
}

/* ===================================================================== */
/* calcLikeDoubleSingle / calcLikeDouble (src/calcLikeDouble.cc)         */
/* ===================================================================== */

typedef struct orc_field2 {
  double *d1, *d2;
  size_t n;
  uint32_t *bins;
  uint32_t nb1, nb2;
  double bc01, bd1, bc02, bd2;
  double sigma_base1, sigma_base2, like_norm, like_offset;
} orc_field2;

typedef struct orc_bset2 {
  orc_dbeam bm;
  orc_field2 *fields;
  unsigned nfields;
  double minR1, maxR1, minR2, maxR2;
  double maxsig1, maxsig2, exp_conf1, exp_conf2;
  double mind1, maxd1, mind2, maxd2;
} orc_bset2;

struct orc_like2 {
  unsigned fftsize, nedge;
  int edge_integ;
  orc_lnmodel model;
  orc_bset2 *sets;
  unsigned nsets;
  int setup_done;
  orc_err err;
  orc_pdfac2 *fac;
};

orc_like2 *orc_like2_new(unsigned fftsize, unsigned nedge, int edge_integ) {
  if (fftsize < 8 || (fftsize & (fftsize - 1)) || nedge < 2) return NULL;
  orc_like2 *L = (orc_like2 *)calloc(1, sizeof(orc_like2));
  L->fftsize = fftsize;
  L->nedge = nedge;
  L->edge_integ = edge_integ;
  orc_ln_init(&L->model);
  return L;
}

static void orc2_dbeam_copy(orc_dbeam *dst, const orc_dbeam *src) {
  memcpy(dst, src, sizeof *dst);
#define DUP(arr, cnt)                                                        \
  do {                                                                       \
    for (int s = 0; s < 4; ++s)                                              \
      if (src->arr[s]) {                                                     \
        dst->arr[s] = (double *)malloc(sizeof(double) * (cnt));              \
        memcpy(dst->arr[s], src->arr[s], sizeof(double) * (cnt));            \
      }                                                                      \
  } while (0)
  DUP(pix1, src->npix[s]);
  DUP(pix2, src->npix[s]);
  DUP(ipix1, src->npix[s]);
  DUP(ipix2, src->npix[s]);
  DUP(logratio, src->npix[s]);
  DUP(binw, src->nhist[s]);
  DUP(binv1, src->nhist[s]);
  DUP(binv2, src->nhist[s]);
  DUP(binlogratio, src->nhist[s]);
#undef DUP
}

void orc_like2_free(orc_like2 *L) {
  if (!L) return;
  for (unsigned b = 0; b < L->nsets; ++b) {
    orc_bset2 *s = &L->sets[b];
    for (unsigned f = 0; f < s->nfields; ++f) {
      free(s->fields[f].d1);
      free(s->fields[f].d2);
      free(s->fields[f].bins);
    }
    free(s->fields);
    orc_dbeam tmp = s->bm;
    for (int k = 0; k < 4; ++k) {
      free(tmp.pix1[k]);
      free(tmp.pix2[k]);
      free(tmp.ipix1[k]);
      free(tmp.ipix2[k]);
      free(tmp.logratio[k]);
      free(tmp.binw[k]);
      free(tmp.binv1[k]);
      free(tmp.binv2[k]);
      free(tmp.binlogratio[k]);
    }
  }
  free(L->sets);
  orc_pdfac2_free(L->fac);
  free(L);
}

const orc_err *orc_like2_error(const orc_like2 *L) { return &L->err; }

int orc_like2_set_positions(orc_like2 *L, unsigned nk, const double *kp,
                            unsigned ns, const double *sp, unsigned no,
                            const double *op, double min_sigma) {
  int rc = orc_ln_set_positions(&L->model, nk, kp, ns, sp, no, op, &L->err);
  if (rc) return rc;
  return orc_ln_set_min_sigma(&L->model, min_sigma, &L->err);
}

int orc_like2_add_beamset(orc_like2 *L, const orc_dbeam *bm) {
  L->sets = (orc_bset2 *)realloc(L->sets, sizeof(orc_bset2) * (L->nsets + 1));
  orc_bset2 *s = &L->sets[L->nsets];
  memset(s, 0, sizeof *s);
  orc2_dbeam_copy(&s->bm, bm);
  s->minR1 = s->maxR1 = s->minR2 = s->maxR2 = NAN;
  return (int)L->nsets++;
}

int orc_like2_add_field(orc_like2 *L, int b, const double *d1,
                        const double *d2, size_t npix, const uint32_t *bins,
                        uint32_t nb1, uint32_t nb2, const double *bininfo,
                        double sigma_base1, double sigma_base2,
                        double like_norm) {
  if (b < 0 || (unsigned)b >= L->nsets || npix == 0) return ORC_E_ARG;
  orc_bset2 *s = &L->sets[b];
  s->fields = (orc_field2 *)realloc(s->fields, sizeof(orc_field2) * (s->nfields + 1));
  orc_field2 *f = &s->fields[s->nfields];
  memset(f, 0, sizeof *f);
  f->n = npix;
  f->d1 = (double *)malloc(sizeof(double) * npix);
  f->d2 = (double *)malloc(sizeof(double) * npix);
  memcpy(f->d1, d1, sizeof(double) * npix);
  memcpy(f->d2, d2, sizeof(double) * npix);
  if (bins) {
    f->nb1 = nb1;
    f->nb2 = nb2;
    f->bins = (uint32_t *)malloc(sizeof(uint32_t) * nb1 * nb2);
    memcpy(f->bins, bins, sizeof(uint32_t) * nb1 * nb2);
    f->bc01 = bininfo[0];
    f->bd1 = bininfo[1];
    f->bc02 = bininfo[2];
    f->bd2 = bininfo[3];
  }
  f->sigma_base1 = sigma_base1;
  f->sigma_base2 = sigma_base2;
  f->like_norm = like_norm;
  return (int)s->nfields++;
}

/* setSigmaBase1/2 (:360-420), setRRange (:496-534) via setRRanges (:1093) */
int orc_like2_finalize_setup(orc_like2 *L, const float *theta, unsigned npar,
                             double exp_conf1, double exp_conf2) {
  int rc = orc_ln_set_params(&L->model, npar, theta, &L->err);
  if (rc) return rc;
  for (unsigned b = 0; b < L->nsets; ++b) {
    orc_bset2 *s = &L->sets[b];
    s->exp_conf1 = exp_conf1;
    s->exp_conf2 = exp_conf2;
    s->maxsig1 = s->fields[0].sigma_base1;
    s->maxsig2 = s->fields[0].sigma_base2;
    for (unsigned f = 1; f < s->nfields; ++f) {
      if (s->fields[f].sigma_base1 > s->maxsig1) s->maxsig1 = s->fields[f].sigma_base1;
      if (s->fields[f].sigma_base2 > s->maxsig2) s->maxsig2 = s->fields[f].sigma_base2;
    }
    double mm[4];
    orc_data2_minmax(s->fields[0].d1, s->fields[0].d2, s->fields[0].n, mm);
    s->mind1 = mm[0];
    s->maxd1 = mm[1];
    s->mind2 = mm[2];
    s->maxd2 = mm[3];
    for (unsigned f = 1; f < s->nfields; ++f) {
      orc_data2_minmax(s->fields[f].d1, s->fields[f].d2, s->fields[f].n, mm);
      if (mm[0] < s->mind1) s->mind1 = mm[0];
      if (mm[1] > s->maxd1) s->maxd1 = mm[1];
      if (mm[2] < s->mind2) s->mind2 = mm[2];
      if (mm[3] > s->maxd2) s->maxd2 = mm[3];
    }
    for (unsigned f = 0; f < s->nfields; ++f) {
      orc_field2 *fd = &s->fields[f];
      double var1 = fd->sigma_base1 * fd->sigma_base1 + exp_conf1 * exp_conf1;
      double var2 = fd->sigma_base2 * fd->sigma_base2 + exp_conf2 * exp_conf2;
      double sigprod = sqrt(var1 * var2);
      fd->like_offset = 0.0;
      if (sigprod > 0) fd->like_offset = -(double)fd->n * (log(orc2_two_pi * sigprod) + 1.0);
    }
    double raw[4];
    rc = orc_ln_rrange(&L->model, &s->bm, raw, &L->err);
    if (rc) return rc;
    const double safetyfac = 1.1;
    double sigma1 = sqrt(exp_conf1 * exp_conf1 + s->maxsig1 * s->maxsig1);
    double sigma2 = sqrt(exp_conf2 * exp_conf2 + s->maxsig2 * s->maxsig2);
    s->minR1 = safetyfac * raw[0];
    s->maxR1 = safetyfac * raw[1] + orc2_n_zero_pad * sigma1;
    s->minR2 = safetyfac * raw[2];
    s->maxR2 = safetyfac * raw[3] + orc2_n_zero_pad * sigma2;
    const orc_dbeam *bm = &s->bm;
    if (s->mind1 < s->minR1 && (bm->hassign[2] || bm->hassign[3])) s->minR1 = s->mind1;
    if (s->maxd1 > s->maxR1 && (bm->hassign[0] || bm->hassign[1])) s->maxR1 = s->maxd1;
    if (s->mind2 < s->minR2 && (bm->hassign[1] || bm->hassign[3])) s->minR2 = s->mind2;
    if (s->maxd2 > s->maxR2 && (bm->hassign[0] || bm->hassign[2])) s->maxR2 = s->maxd2;
  }
  L->setup_done = 1;
  return ORC_OK;
}

int orc_like2_set_rrange(orc_like2 *L, int b, const double *r) {
  if (b < 0 || (unsigned)b >= L->nsets) return ORC_E_ARG;
  orc_bset2 *s = &L->sets[b];
  s->minR1 = r[0];
  s->maxR1 = r[1];
  s->minR2 = r[2];
  s->maxR2 = r[3];
  return ORC_OK;
}

int orc_like2_beamset_info(const orc_like2 *L, int b, double *out) {
  if (b < 0 || (unsigned)b >= L->nsets) return ORC_E_ARG;
  const orc_bset2 *s = &L->sets[b];
  out[0] = s->minR1;
  out[1] = s->maxR1;
  out[2] = s->minR2;
  out[3] = s->maxR2;
  out[4] = s->maxsig1;
  out[5] = s->maxsig2;
  out[6] = s->mind1;
  out[7] = s->maxd1;
  out[8] = s->mind2;
  out[9] = s->maxd2;
  out[10] = s->nfields;
  out[11] = s->nfields ? s->fields[0].like_offset : NAN;
  return ORC_OK;
}

/* calcLikeDoubleSingle::getLogLike (:558-620) for one beam set.
   A model that fails isValid() is reported as ORC_E_INVALID_MODEL (the
   reference would carry NaN R values into the transforms). */
static int orc2_bset_loglike(const orc_like2 *L, const orc_bset2 *s,
                             orc_lnmodel *model, orc_pdfac2 *p, double sm1,
                             double sm2, double *ll, orc_err *err) {
  if (s->maxsig1 * sm1 < 0.0 || s->maxsig2 * sm2 < 0.0) {
    *ll = 1e25; /* bad_like */
    return ORC_E_BADLIKE;
  }
  if (!orc_ln_is_valid(model))
    return orc2_fail(err, ORC_E_INVALID_MODEL, "PDFactoryDouble", "initPD", "Invalid model");
  int rc = orc2_initpd(p, s->minR1, s->maxR1, s->minR2, s->maxR2, model, &s->bm,
                       L->edge_integ, err);
  if (rc) return rc;
  double tot = 0.0;
  for (unsigned f = 0; f < s->nfields; ++f) {
    const orc_field2 *fd = &s->fields[f];
    rc = orc2_getpd(p, sm1 * fd->sigma_base1, sm2 * fd->sigma_base2, 1, err);
    if (rc) return rc;
    const unsigned n = p->n;
    double curr;
    if (fd->bins)
      curr = orc_pd2_loglike_binned(n, n, p->pd, p->pd_minflux1, p->dflux1, p->pd_minflux2,
                                    p->dflux2, 1, fd->bins, fd->nb1, fd->nb2, fd->bc01,
                                    fd->bd1, fd->bc02, fd->bd2);
    else
      curr = orc_pd2_loglike_unbinned(n, n, p->pd, p->pd_minflux1, p->dflux1, p->pd_minflux2,
                                      p->dflux2, 1, fd->d1, fd->d2, fd->n);
    tot += (curr - fd->like_offset) / fd->like_norm;
  }
  *ll = tot;
  return ORC_OK;
}

/* calcLikeDouble::getLogLike (:1195-1226), priors not included */
static int orc2_loglike_ws(const orc_like2 *L, orc_lnmodel *model, orc_pdfac2 *p,
                           const float *theta, unsigned npar, double *out,
                           orc_err *err) {
  const unsigned nmp = model->nknots + model->nsigma + model->noffset;
  if (npar < nmp + 2)
    return orc2_fail(err, ORC_E_ARG, "calcLikeDouble", "getLogLike", "Not enough elements in params");
  int rc = orc_ln_set_params(model, npar, theta, err);
  if (rc) return rc;
  double sm1 = (double)theta[nmp], sm2 = (double)theta[nmp + 1];
  double total = 0.0;
  int bad = 0;
  for (unsigned b = 0; b < L->nsets; ++b) {
    double v = 0.0;
    rc = orc2_bset_loglike(L, &L->sets[b], model, p, sm1, sm2, &v, err);
    if (rc == ORC_E_BADLIKE) {
      bad = 1;
      total += v;
      continue;
    }
    if (rc) return rc;
    total += v;
  }
  *out = total;
  return bad ? ORC_E_BADLIKE : ORC_OK;
}

int orc_like2_loglike(orc_like2 *L, const float *theta, unsigned npar,
                      double *loglike) {
  if (!L->setup_done) return ORC_E_ARG;
  if (!L->fac) L->fac = orc_pdfac2_new(L->fftsize, L->nedge);
  orc_lnmodel model = L->model;
  return orc2_loglike_ws(L, &model, L->fac, theta, npar, loglike, &L->err);
}

typedef struct {
  orc_like2 *L;
  const float *theta;
  unsigned nwalk, stride, npar, tid, nthreads;
  double *loglike;
  int32_t *status;
} orc2_job;

static void *orc2_worker(void *arg) {
  orc2_job *j = (orc2_job *)arg;
  orc_pdfac2 *p = orc_pdfac2_new(j->L->fftsize, j->L->nedge);
  orc_lnmodel model = j->L->model;
  orc_err err;
  for (unsigned w = j->tid; w < j->nwalk; w += j->nthreads) {
    memset(&err, 0, sizeof err);
    double v = NAN;
    int rc = orc2_loglike_ws(j->L, &model, p, j->theta + (size_t)w * j->stride, j->npar, &v, &err);
    j->loglike[w] = (rc && rc != ORC_E_BADLIKE) ? NAN : v;
    j->status[w] = rc;
  }
  orc_pdfac2_free(p);
  return NULL;
}

int orc_like2_loglike_batch(orc_like2 *L, const float *theta, unsigned nwalk,
                            unsigned stride, unsigned npar, double *loglike,
                            int32_t *status, unsigned nthreads) {
  if (!L->setup_done) return ORC_E_ARG;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 64) nthreads = 64;
  pthread_t th[64];
  orc2_job jobs[64];
  for (unsigned t = 0; t < nthreads; ++t) {
    jobs[t] = (orc2_job){L, theta, nwalk, stride, npar, t, nthreads, loglike, status};
    if (nthreads == 1)
      orc2_worker(&jobs[t]);
    else
      pthread_create(&th[t], NULL, orc2_worker, &jobs[t]);
  }
  if (nthreads > 1)
    for (unsigned t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  return ORC_OK;
}

int orc_like2_initpd(orc_like2 *L, int b, const float *theta, unsigned npar,
                     double *rvals, double *prtrans, double *scalars) {
  if (!L->setup_done || b < 0 || (unsigned)b >= L->nsets) return ORC_E_ARG;
  if (!L->fac) L->fac = orc_pdfac2_new(L->fftsize, L->nedge);
  int rc = orc_ln_set_params(&L->model, npar, theta, &L->err);
  if (rc) return rc;
  const orc_bset2 *s = &L->sets[b];
  orc_pdfac2 *p = L->fac;
  rc = orc2_initpd(p, s->minR1, s->maxR1, s->minR2, s->maxR2, &L->model, &s->bm,
                   L->edge_integ, &L->err);
  if (rc) return rc;
  const size_t n = L->fftsize;
  /* rvals was transformed in place only in a copy: orc_rfft2 reads rvals */
  if (rvals) memcpy(rvals, p->rvals, sizeof(double) * n * n);
  if (prtrans) memcpy(prtrans, p->prtrans, sizeof(double) * 2 * n * (n / 2 + 1));
  if (scalars) {
    scalars[0] = p->dflux1;
    scalars[1] = p->dflux2;
    scalars[2] = p->mn1;
    scalars[3] = p->mn2;
    scalars[4] = p->var1;
    scalars[5] = p->var2;
    scalars[6] = p->shift1;
    scalars[7] = p->shift2;
    scalars[8] = p->doshift1;
    scalars[9] = p->doshift2;
    scalars[10] = p->r0;
    scalars[11] = p->wrap1;
  }
  return ORC_OK;
}

int orc_like2_getpd(orc_like2 *L, int b, int f, double sigmult1,
                    double sigmult2, int setlog, double *pd_out, double *info) {
  if (!L->fac || b < 0 || (unsigned)b >= L->nsets) return ORC_E_ARG;
  const orc_bset2 *s = &L->sets[b];
  if (f < 0 || (unsigned)f >= s->nfields) return ORC_E_ARG;
  orc_pdfac2 *p = L->fac;
  int rc = orc2_getpd(p, sigmult1 * s->fields[f].sigma_base1,
                      sigmult2 * s->fields[f].sigma_base2, setlog, &L->err);
  if (rc) return rc;
  const size_t n = p->n;
  if (pd_out) memcpy(pd_out, p->pd, sizeof(double) * n * n);
  if (info) {
    info[0] = p->pd_minflux1;
    info[1] = p->pd_minflux2;
    info[2] = p->dflux1;
    info[3] = p->dflux2;
    info[4] = p->split1;
    info[5] = p->split2;
    info[6] = p->sg1pos;
    info[7] = p->sg2pos;
  }
  return ORC_OK;
}
