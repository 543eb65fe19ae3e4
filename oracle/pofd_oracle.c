/*
 * pofd_oracle.c -- CPU restatement of the reference 1D P(D) likelihood path.
 *
 * TEST INFRASTRUCTURE ONLY (see pofd_oracle.h for the rules and the pins).
 * Each function cites the reference file:line it follows.  Paths are
 * relative to the reference repository root.
 */
#define _GNU_SOURCE
#include "pofd_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* include/global_settings.h:20-21, 66-74 */
#define ORC_PI 3.141592653589793238462643383279502884197
static const double orc_two_pi = 2.0 * ORC_PI;
static const double orc_n_zero_pad = 7.5;
static const double orc_smalllogval = -30.;
/* logfac = log2(10.0), log2toe = log(2.) are evaluated at run time exactly
   as the reference does (global_settings.h:69,73) */
static double orc_logfac(void) { return log2(10.0); }
static double orc_log2toe(void) { return log(2.); }

static int orc_fail(orc_err *e, int code, const char *cls, const char *method,
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
/* GSL natural cubic spline, restated (GSL interp/cspline.c,            */
/* linalg/tridiag.c).  Called from src/PDFactory.cc:482,522 and         */
/* src/numberCountsKnotsSpline.cc:271 via gsl_spline_init.               */
/* ===================================================================== */

/* gsl_linalg_solve_symm_tridiag: LDL^T of the symmetric tridiagonal
   matrix (diag, offdiag), forward substitution, scaling, back
   substitution, in exactly that order. */
static void orc_solve_symm_tridiag(size_t N, const double *diag,
                                   const double *offdiag, const double *b,
                                   double *x) {
  double *gamma = (double *)malloc(sizeof(double) * N);
  double *alpha = (double *)malloc(sizeof(double) * N);
  double *c = (double *)malloc(sizeof(double) * N);
  double *z = (double *)malloc(sizeof(double) * N);
  size_t i, j;
  alpha[0] = diag[0];
  gamma[0] = offdiag[0] / alpha[0];
  for (i = 1; i < N - 1; i++) {
    alpha[i] = diag[i] - offdiag[i - 1] * gamma[i - 1];
    gamma[i] = offdiag[i] / alpha[i];
  }
  if (N > 1) alpha[N - 1] = diag[N - 1] - offdiag[N - 2] * gamma[N - 2];
  z[0] = b[0];
  for (i = 1; i < N; i++) z[i] = b[i] - gamma[i - 1] * z[i - 1];
  for (i = 0; i < N; i++) c[i] = z[i] / alpha[i];
  x[N - 1] = c[N - 1];
  if (N >= 2)
    for (i = N - 2, j = 0; j <= N - 2; j++, i--) x[i] = c[i] - gamma[i] * x[i + 1];
  free(gamma);
  free(alpha);
  free(c);
  free(z);
}

/* cspline_init with natural boundary conditions */
int orc_cspline_init(size_t n, const double *xa, const double *ya,
                     double *c_out) {
  size_t i;
  size_t max_index = n - 1;
  size_t sys_size = max_index - 1;
  if (n < 2) return ORC_E_ARG;
  for (i = 1; i < n; ++i)
    if (!(xa[i] > xa[i - 1])) return ORC_E_ARG; /* gsl_spline_init EINVAL */
  c_out[0] = 0.0;
  c_out[max_index] = 0.0;
  if (n == 2) return ORC_OK;
  double *offdiag = (double *)malloc(sizeof(double) * sys_size);
  double *diag = (double *)malloc(sizeof(double) * sys_size);
  double *g = (double *)malloc(sizeof(double) * sys_size);
  for (i = 0; i < sys_size; i++) {
    const double h_i = xa[i + 1] - xa[i];
    const double h_ip1 = xa[i + 2] - xa[i + 1];
    const double ydiff_i = ya[i + 1] - ya[i];
    const double ydiff_ip1 = ya[i + 2] - ya[i + 1];
    const double g_i = (h_i != 0.0) ? 1.0 / h_i : 0.0;
    const double g_ip1 = (h_ip1 != 0.0) ? 1.0 / h_ip1 : 0.0;
    offdiag[i] = h_ip1;
    diag[i] = 2.0 * (h_ip1 + h_i);
    g[i] = 3.0 * (ydiff_ip1 * g_ip1 - ydiff_i * g_i);
  }
  if (sys_size == 1)
    c_out[1] = g[0] / diag[0];
  else
    orc_solve_symm_tridiag(sys_size, diag, offdiag, g, c_out + 1);
  free(offdiag);
  free(diag);
  free(g);
  return ORC_OK;
}

/* gsl_interp_bsearch */
size_t orc_bsearch(const double *x, double xv, size_t lo, size_t hi) {
  size_t ilo = lo, ihi = hi;
  while (ihi > ilo + 1) {
    size_t i = (ihi + ilo) / 2;
    if (x[i] > xv)
      ihi = i;
    else
      ilo = i;
  }
  return ilo;
}

/* cspline_eval + coeff_calc.  Points outside [x0, x_{n-1}] (which the
   reference never produces except by rounding at the fill edges) evaluate
   the edge cubic rather than raising GSL_EDOM. */
double orc_cspline_eval(size_t n, const double *xa, const double *ya,
                        const double *c, double xv) {
  size_t index;
  if (xv <= xa[0])
    index = 0;
  else if (xv >= xa[n - 1])
    index = n - 2;
  else
    index = orc_bsearch(xa, xv, 0, n - 1);
  const double x_hi = xa[index + 1];
  const double x_lo = xa[index];
  const double dx = x_hi - x_lo;
  if (dx > 0.0) {
    const double y_lo = ya[index];
    const double y_hi = ya[index + 1];
    const double dy = y_hi - y_lo;
    const double delx = xv - x_lo;
    const double c_i = c[index];
    const double c_ip1 = c[index + 1];
    const double b_i = (dy / dx) - dx * (c_ip1 + 2.0 * c_i) / 3.0;
    const double d_i = (c_ip1 - c_i) / (3.0 * dx);
    return y_lo + delx * (b_i + delx * (c_i + delx * d_i));
  }
  return 0.0;
}

/* ===================================================================== */
/* beam: src/beam.cc:295-458 (readFile after FITS), 467-556 (histogram) */
/* ===================================================================== */

orc_beam *orc_beam_new(void) {
  orc_beam *b = (orc_beam *)calloc(1, sizeof(orc_beam));
  return b;
}

static void orc_beam_clear(orc_beam *b) {
  free(b->pospix);
  free(b->negpix);
  free(b->posinv);
  free(b->neginv);
  free(b->posweights);
  free(b->poshistval);
  free(b->negweights);
  free(b->neghistval);
  memset(b, 0, sizeof(*b));
}

void orc_beam_free(orc_beam *b) {
  if (!b) return;
  orc_beam_clear(b);
  free(b);
}

static int orc_cmp_double(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

int orc_beam_set(orc_beam *bm, const double *pixarr, size_t n, double pixsize,
                 double MINVAL, orc_err *err) {
  orc_beam_clear(bm);
  bm->pixsize = pixsize;
  /* src/beam.cc:307-310 */
  bm->minval = (MINVAL <= 0.0) ? 0.0 : MINVAL;
  double minval = bm->minval, negminval = -minval;
  /* src/beam.cc:383-388 */
  uint32_t npos = 0, nneg = 0;
  for (size_t i = 0; i < n; ++i) {
    if (isnan(pixarr[i]) || isinf(pixarr[i])) continue;
    if (pixarr[i] > minval) npos++;
    if (pixarr[i] < negminval) nneg++;
  }
  if (npos == 0 && nneg == 0)
    return orc_fail(err, ORC_E_ARG, "beam", "readFile",
                    "Beam file has only zero values");
  /* src/beam.cc:395-416 */
  if (npos > 0) {
    uint32_t ctr = 0;
    bm->pospix = (double *)malloc(sizeof(double) * npos);
    for (size_t i = 0; i < n; ++i) {
      if (isnan(pixarr[i]) || isinf(pixarr[i])) continue;
      if (pixarr[i] > minval) bm->pospix[ctr++] = pixarr[i];
    }
    bm->haspos = 1;
  }
  if (nneg > 0) {
    uint32_t ctr = 0;
    bm->negpix = (double *)malloc(sizeof(double) * nneg);
    for (size_t i = 0; i < n; ++i) {
      if (isnan(pixarr[i]) || isinf(pixarr[i])) continue;
      if (pixarr[i] < negminval) bm->negpix[ctr++] = fabs(pixarr[i]);
    }
    bm->hasneg = 1;
  }
  bm->npos = npos;
  bm->nneg = nneg;
  /* src/beam.cc:419-421 */
  if (bm->haspos) qsort(bm->pospix, npos, sizeof(double), orc_cmp_double);
  if (bm->hasneg) qsort(bm->negpix, nneg, sizeof(double), orc_cmp_double);
  /* src/beam.cc:423-444 */
  if (bm->haspos) {
    bm->totpos = bm->pospix[0];
    bm->totpossq = bm->totpos * bm->totpos;
    for (uint32_t i = 1; i < npos; ++i) {
      double val = bm->pospix[i];
      bm->totpos += val;
      bm->totpossq += val * val;
    }
  }
  if (bm->hasneg) {
    bm->totneg = bm->negpix[0];
    bm->totnegsq = bm->totneg * bm->totneg;
    for (uint32_t i = 1; i < nneg; ++i) {
      double val = bm->negpix[i];
      bm->totneg += val;
      bm->totnegsq += val * val;
    }
  }
  /* src/beam.cc:446-456 */
  if (bm->haspos) {
    bm->posinv = (double *)malloc(sizeof(double) * npos);
    for (uint32_t i = 0; i < npos; ++i) bm->posinv[i] = 1.0 / bm->pospix[i];
  }
  if (bm->hasneg) {
    bm->neginv = (double *)malloc(sizeof(double) * nneg);
    for (uint32_t i = 0; i < nneg; ++i) bm->neginv[i] = 1.0 / bm->negpix[i];
  }
  return ORC_OK;
}

/* one sign of beam::makeHistogram, src/beam.cc:493-524 */
static void orc_hist_one(const double *pix, uint32_t npix, uint32_t nbins,
                         uint32_t *nout, double **wout, double **vout) {
  const double log2outscale = 0.0014419741739063218; /* log2(1.001) */
  uint32_t *tmp_wt = (uint32_t *)calloc(nbins, sizeof(uint32_t));
  double *tmp_val = (double *)calloc(nbins, sizeof(double));
  double minlog2val = log2(pix[0]) - log2outscale;
  double maxlog2val = log2(pix[npix - 1]) + log2outscale;
  double histstep = (maxlog2val - minlog2val) / (double)nbins;
  for (uint32_t i = 0; i < npix; ++i) {
    double curr_val = pix[i];
    uint32_t idx = (uint32_t)((log2(curr_val) - minlog2val) / histstep);
    tmp_wt[idx] += 1;
    tmp_val[idx] += curr_val;
  }
  uint32_t nb = 0;
  for (uint32_t i = 0; i < nbins; ++i)
    if (tmp_wt[i] > 0) ++nb;
  double *w = (double *)malloc(sizeof(double) * (nb ? nb : 1));
  double *v = (double *)malloc(sizeof(double) * (nb ? nb : 1));
  uint32_t idx = 0;
  for (uint32_t i = 0; i < nbins; ++i)
    if (tmp_wt[i] > 0) {
      w[idx] = (double)tmp_wt[i];
      v[idx] = tmp_val[i] / w[idx];
      ++idx;
    }
  for (uint32_t i = 0; i < nb; ++i) v[i] = 1.0 / v[i];
  free(tmp_wt);
  free(tmp_val);
  *nout = nb;
  *wout = w;
  *vout = v;
}

int orc_beam_make_histogram(orc_beam *bm, uint32_t NBINS, orc_err *err) {
  /* src/beam.cc:467-491; histothresh = 20 (include/beam.h:31) */
  if (NBINS == 0)
    return orc_fail(err, ORC_E_ARG, "beam", "makeHistogram",
                    "Invalid (non-positive) nbins");
  bm->nbins = NBINS;
  bm->is_pos_hist = bm->is_neg_hist = 0;
  bm->posnbins = bm->negnbins = 0;
  free(bm->posweights);
  free(bm->poshistval);
  free(bm->negweights);
  free(bm->neghistval);
  bm->posweights = bm->poshistval = bm->negweights = bm->neghistval = NULL;
  int do_pos = bm->haspos && (bm->npos >= 20);
  int do_neg = bm->hasneg && (bm->nneg >= 20);
  if (do_pos) {
    orc_hist_one(bm->pospix, bm->npos, NBINS, &bm->posnbins, &bm->posweights,
                 &bm->poshistval);
    bm->is_pos_hist = 1;
  }
  if (do_neg) {
    orc_hist_one(bm->negpix, bm->nneg, NBINS, &bm->negnbins, &bm->negweights,
                 &bm->neghistval);
    bm->is_neg_hist = 1;
  }
  return ORC_OK;
}

static int orc_beam_isneghist(const orc_beam *bm) {
  /* include/beam.h:97 returns is_pos_histogrammed (reproduced quirk) */
  return bm->is_pos_hist;
}

void orc_beam_info(const orc_beam *bm, double *out) {
  double conv = bm->pixsize / 3600.0;
  memset(out, 0, 16 * sizeof(double));
  out[0] = bm->npos;
  out[1] = bm->nneg;
  out[2] = bm->haspos;
  out[3] = bm->hasneg;
  out[4] = bm->pixsize;
  out[5] = bm->minval;
  out[6] = bm->is_pos_hist;
  out[7] = bm->is_neg_hist;
  out[8] = bm->posnbins;
  out[9] = bm->negnbins;
  /* getEffectiveArea* src/beam.cc:614-668 */
  double area = 0.0, areapix = 0.0, areasq = 0.0;
  if (bm->haspos) {
    area = bm->totpos;
    areapix = bm->totpos;
    areasq = bm->totpossq * conv * conv;
  }
  if (bm->hasneg) {
    area += bm->totneg;
    areapix += bm->totneg;
    areasq += bm->totnegsq * conv * conv;
  }
  out[10] = area * conv * conv;
  out[11] = bm->haspos ? bm->totpos * conv * conv : 0.0;
  out[12] = bm->hasneg ? bm->totneg * conv * conv : 0.0;
  out[13] = areapix;
  out[14] = areasq;
  out[15] = bm->haspos ? bm->pospix[bm->npos - 1] : NAN;
}

const double *orc_beam_array(const orc_beam *bm, int which) {
  switch (which) {
  case 0: return bm->pospix;
  case 1: return bm->negpix;
  case 2: return bm->posinv;
  case 3: return bm->neginv;
  case 4: return bm->posweights;
  case 5: return bm->poshistval;
  case 6: return bm->negweights;
  case 7: return bm->neghistval;
  }
  return NULL;
}

static void orc_beam_copy(orc_beam *dst, const orc_beam *src) {
  *dst = *src;
#define DUP(f, cnt)                                                        \
  if (src->f) {                                                            \
    dst->f = (double *)malloc(sizeof(double) * ((cnt) ? (cnt) : 1));       \
    memcpy(dst->f, src->f, sizeof(double) * (cnt));                        \
  }
  DUP(pospix, src->npos);
  DUP(negpix, src->nneg);
  DUP(posinv, src->npos);
  DUP(neginv, src->nneg);
  DUP(posweights, src->posnbins);
  DUP(poshistval, src->posnbins);
  DUP(negweights, src->negnbins);
  DUP(neghistval, src->negnbins);
#undef DUP
}

/* ===================================================================== */
/* knot model: src/numberCountsKnots.cc, src/numberCountsKnotsSpline.cc  */
/* ===================================================================== */

int orc_knots_set_positions(orc_knots *m, unsigned n, const double *S,
                            orc_err *err) {
  /* src/numberCountsKnots.cc:218-238 + numberCountsKnotsSpline.cc:238-244 */
  if (n < 2 || n > ORC_MAXKNOTS)
    return orc_fail(err, ORC_E_ARG, "numberCountsKnots", "setKnotPositions",
                    "Unsupported number of knots %u", n);
  for (unsigned i = 0; i < n; ++i)
    if (S[i] <= 0.0)
      return orc_fail(err, ORC_E_ARG, "numberCountsKnots", "setKnots",
                      "Negative knot positions not allowed");
  for (unsigned i = 1; i < n; ++i)
    if (S[i] <= S[i - 1])
      return orc_fail(err, ORC_E_ARG, "numberCountsKnots", "setKnotPositions",
                      "Knot positions not monotonically increasing");
  m->nknots = n;
  m->loaded = 0;
  for (unsigned i = 0; i < n; ++i) {
    m->knots[i] = S[i];
    m->logknots[i] = log2(m->knots[i]);
  }
  return ORC_OK;
}

int orc_knots_set_params(orc_knots *m, unsigned np, const float *F,
                         orc_err *err) {
  /* src/numberCountsKnotsSpline.cc:265-274 */
  if (m->nknots > np)
    return orc_fail(err, ORC_E_ARG, "numberCountsKnotsSpline", "setKnots",
                    "Number of knot values different than expected");
  double lf = orc_logfac();
  for (unsigned i = 0; i < m->nknots; ++i)
    m->logknotvals[i] = lf * (double)F[i];
  /* gsl_spline_init on (logknots, logknotvals): GSL's own validation
     rejects NaN knot values only through the error handler; a NaN value
     makes isValid() false below, which initPD reports first. */
  orc_cspline_init(m->nknots, m->logknots, m->logknotvals, m->c);
  m->loaded = 1;
  return ORC_OK;
}

int orc_knots_is_valid(const orc_knots *m) {
  /* src/numberCountsKnots.cc:273-284 */
  if (m->nknots == 0) return 0;
  if (!m->loaded) return 0;
  for (unsigned i = 0; i < m->nknots; ++i)
    if (isnan(m->knots[i])) return 0;
  if (m->knots[0] <= 0.0) return 0;
  for (unsigned i = 1; i < m->nknots; ++i)
    if (m->knots[i] <= m->knots[i - 1]) return 0;
  for (unsigned i = 0; i < m->nknots; ++i)
    if (isnan(m->logknotvals[i])) return 0;
  return 1;
}

static double orc_knots_spline(const orc_knots *m, double lx) {
  return orc_cspline_eval(m->nknots, m->logknots, m->logknotvals, m->c, lx);
}

double orc_knots_number_counts(const orc_knots *m, double fdens) {
  /* src/numberCountsKnotsSpline.cc:284-289 */
  if (m->nknots < 2) return NAN;
  if (!orc_knots_is_valid(m)) return NAN;
  if (fdens < m->knots[0] || fdens >= m->knots[m->nknots - 1]) return 0.0;
  return exp2(orc_knots_spline(m, log2(fdens)));
}

void orc_knots_rrange(const orc_knots *m, const orc_beam *bm, double *minRF,
                      double *maxRF) {
  /* src/numberCountsKnots.cc:345-365 */
  double maxknot = m->knots[m->nknots - 1];
  if (bm->haspos) {
    *maxRF = maxknot * bm->pospix[bm->npos - 1];
    *minRF = bm->hasneg ? -maxknot * bm->negpix[bm->nneg - 1] : 0.0;
  } else {
    *maxRF = 0.0;
    *minRF = -maxknot * bm->negpix[bm->nneg - 1];
  }
}

double orc_knots_getR1(const orc_knots *m, double fluxdensity,
                       const orc_beam *bm) {
  /* scalar getR, src/numberCountsKnotsSpline.cc:352-429 */
  if (!orc_knots_is_valid(m)) return NAN;
  if (fluxdensity > 0 && !bm->haspos) return 0.0;
  else if (fluxdensity < 0 && !bm->hasneg) return 0.0;
  else if (fluxdensity == 0) return 0.0;
  double minknot = m->knots[0], maxknot = m->knots[m->nknots - 1];
  double rmin, rmax;
  orc_knots_rrange(m, bm, &rmin, &rmax);
  if (fluxdensity >= rmax) return 0.0;
  if (fluxdensity <= rmin) return 0.0;
  double prefac = bm->pixsize / 3600.0;
  prefac = prefac * prefac;
  double retval = 0.0;
  if (fluxdensity > 0) {
    if (bm->is_pos_hist) {
      for (uint32_t i = 0; i < bm->posnbins; ++i) {
        double ieta = bm->poshistval[i];
        double currarg = fluxdensity * ieta;
        if (currarg < minknot || currarg >= maxknot) continue;
        double cts = exp2(orc_knots_spline(m, log2(currarg)));
        retval += bm->posweights[i] * ieta * cts;
      }
    } else {
      for (uint32_t i = 0; i < bm->npos; ++i) {
        double ieta = bm->posinv[i];
        double currarg = fluxdensity * ieta;
        if (currarg < minknot || currarg >= maxknot) continue;
        retval += exp2(orc_knots_spline(m, log2(currarg))) * ieta;
      }
    }
  } else if (fluxdensity < 0) {
    if (orc_beam_isneghist(bm)) {
      for (uint32_t i = 0; i < bm->negnbins; ++i) {
        double ieta = bm->neghistval[i];
        double currarg = -fluxdensity * ieta;
        if (currarg < minknot || currarg >= maxknot) continue;
        double cts = exp2(orc_knots_spline(m, log2(currarg)));
        retval += bm->negweights[i] * ieta * cts;
      }
    } else {
      for (uint32_t i = 0; i < bm->nneg; ++i) {
        double ieta = bm->neginv[i];
        double currarg = -fluxdensity * ieta;
        if (currarg < minknot || currarg >= maxknot) continue;
        retval += exp2(orc_knots_spline(m, log2(currarg))) * ieta;
      }
    }
  }
  return prefac * retval;
}

void orc_knots_getR(const orc_knots *m, unsigned n, const double *flux,
                    const orc_beam *bm, double *R) {
  /* array getR, src/numberCountsKnotsSpline.cc:436-529 */
  if (n == 0) return;
  if (!orc_knots_is_valid(m)) {
    for (unsigned i = 0; i < n; ++i) R[i] = NAN;
    return;
  }
  double rmin, rmax;
  orc_knots_rrange(m, bm, &rmin, &rmax);
  double minknot = m->knots[0], maxknot = m->knots[m->nknots - 1];
  double prefac = bm->pixsize / 3600.0;
  prefac = prefac * prefac;
  uint32_t npos = 0, nneg = 0;
  int haspos = bm->haspos, hasneg = bm->hasneg;
  int poshist = 0, neghist = 0;
  const double *wpos = NULL, *wneg = NULL, *ipos = NULL, *ineg = NULL;
  if (haspos) {
    poshist = bm->is_pos_hist;
    npos = poshist ? bm->posnbins : bm->npos;
    if (poshist) wpos = bm->posweights;
    ipos = poshist ? bm->poshistval : bm->posinv;
  }
  if (hasneg) {
    neghist = orc_beam_isneghist(bm);
    nneg = neghist ? bm->negnbins : bm->nneg;
    if (neghist) wneg = bm->negweights;
    ineg = neghist ? bm->neghistval : bm->neginv;
  }
  for (unsigned i = 0; i < n; ++i) {
    double fluxdensity = flux[i];
    double workval = 0.0;
    if (fluxdensity > 0 && haspos && fluxdensity <= rmax) {
      if (poshist) {
        for (uint32_t j = 0; j < npos; ++j) {
          double ieta = ipos[j];
          double currarg = fluxdensity * ieta;
          if (currarg < minknot || currarg >= maxknot) continue;
          double cts = exp2(orc_knots_spline(m, log2(currarg)));
          if (cts > 0) workval += wpos[j] * cts * ieta;
        }
      } else {
        for (uint32_t j = 0; j < npos; ++j) {
          double ieta = ipos[j];
          double currarg = fluxdensity * ieta;
          if (currarg < minknot || currarg >= maxknot) continue;
          double cts = exp2(orc_knots_spline(m, log2(currarg)));
          if (cts > 0) workval += cts * ieta;
        }
      }
    } else if (hasneg && fluxdensity < 0 && fluxdensity >= rmin) {
      if (neghist) {
        for (uint32_t j = 0; j < nneg; ++j) {
          double ieta = ineg[j];
          double currarg = -fluxdensity * ieta;
          if (currarg < minknot || currarg >= maxknot) continue;
          double cts = exp2(orc_knots_spline(m, log2(currarg)));
          if (cts > 0) workval += wneg[j] * cts * ieta;
        }
      } else {
        for (uint32_t j = 0; j < nneg; ++j) {
          double ieta = ineg[j];
          double currarg = -fluxdensity * ieta;
          if (currarg < minknot || currarg >= maxknot) continue;
          double cts = exp2(orc_knots_spline(m, log2(currarg)));
          if (cts > 0) workval += cts * ieta;
        }
      }
    }
    R[i] = prefac * workval;
  }
}

/* Gauss-Legendre nodes and weights on [-1, 1]: Newton iteration on the
   Legendre polynomial (three-term recurrence) in long double */
void orc_gl_nodes(unsigned n, double *x, double *w) {
  const long double pi = 3.141592653589793238462643383279502884L;
  for (unsigned i = 0; i < n; ++i) {
    long double z = cosl(pi * ((long double)i + 0.75L) / ((long double)n + 0.5L));
    long double dp = 0.0L;
    for (int it = 0; it < 100; ++it) {
      long double p0 = 1.0L, p1 = z;
      for (unsigned k = 2; k <= n; ++k) {
        const long double p2 = ((2.0L * k - 1.0L) * z * p1 - (k - 1.0L) * p0) / (long double)k;
        p0 = p1;
        p1 = p2;
      }
      dp = (long double)n * (z * p1 - p0) / (z * z - 1.0L);
      const long double dz = p1 / dp;
      z -= dz;
      if (fabsl(dz) < 1e-19L) break;
    }
    x[i] = (double)z;
    w[i] = (double)(2.0L / ((1.0L - z * z) * dp * dp));
  }
}

/* numberCountsKnotsSpline::splineInt, src/numberCountsKnotsSpline.cc:
   300-321: int S^alpha n(S) dS over [minknot, maxknot).  The reference
   integrates with GSL QAG-41 to relative 1e-5; here the integral is taken
   to full double precision as a 32-point Gauss-Legendre sum in log2 S over
   each knot interval, where the integrand 2^(spline + (alpha+1) v) ln 2 is
   analytic. */
double orc_knots_splineint(const orc_knots *m, double alpha) {
  if (m->nknots < 2 || !orc_knots_is_valid(m)) return NAN;
  enum { NG = 32 };
  double gx[NG], gw[NG];
  orc_gl_nodes(NG, gx, gw);
  long double acc = 0.0L;
  for (unsigned k = 0; k + 1 < m->nknots; ++k) {
    const double v0 = m->logknots[k], v1 = m->logknots[k + 1];
    const double h = 0.5 * (v1 - v0), c = 0.5 * (v1 + v0);
    for (unsigned g = 0; g < NG; ++g) {
      const double v = c + h * gx[g];
      acc += (long double)(gw[g] * h) *
             (long double)exp2(orc_knots_spline(m, v) + (alpha + 1.0) * v);
    }
  }
  return (double)(acc * 0.693147180559945309417232121458L);
}

double orc_knots_regularize(const orc_knots *m, double alpha) {
  /* src/numberCountsKnotsSpline.cc:540-559 */
  if (!m->loaded) return NAN;
  if (m->nknots == 0 || m->nknots == 1) return 0.0;
  double log_penalty = 0.0, mean_deriv = 0.0;
  for (unsigned i = 1; i < m->nknots; ++i) {
    double dp = m->logknots[i] - m->logknots[i - 1];
    double dv = m->logknotvals[i] - m->logknotvals[i - 1];
    mean_deriv += dv / dp;
  }
  mean_deriv /= (double)(m->nknots - 1);
  for (unsigned i = 1; i < m->nknots; ++i) {
    double dp = m->logknots[i] - m->logknots[i - 1];
    double dv = m->logknotvals[i] - m->logknotvals[i - 1];
    double delta = dv / dp - mean_deriv;
    log_penalty -= delta * delta;
  }
  return alpha * log_penalty;
}

/* ===================================================================== */
/* data: src/fitsData.cc:348-437                                          */
/* ===================================================================== */

void orc_data_minmax(const double *d, size_t n, double *mn, double *mx) {
  /* fitsData::getMinMax (unbinned branch) src/fitsData.cc:348-368 */
  double lo = d[0], hi = d[0];
  for (size_t i = 1; i < n; ++i) {
    double c = d[i];
    if (c < lo) lo = c;
    if (c > hi) hi = c;
  }
  *mn = lo;
  *mx = hi;
}

double orc_data_mean_subtract(double *d, size_t n) {
  /* fitsData::getMean + meanSubtract src/fitsData.cc:373-401 */
  double mnval = d[0];
  for (size_t i = 1; i < n; ++i) mnval += d[i];
  mnval /= (double)n;
  for (size_t i = 0; i < n; ++i) d[i] -= mnval;
  return mnval;
}

int orc_data_bin(const double *d, size_t n, uint32_t nbins, uint32_t *bins,
                 double *bincent0, double *bindelta) {
  /* fitsData::applyBinning src/fitsData.cc:406-437 */
  if (n == 0) return ORC_OK;
  if (nbins == 0) return ORC_E_ARG;
  memset(bins, 0, nbins * sizeof(uint32_t));
  double mn, mx;
  orc_data_minmax(d, n, &mn, &mx);
  double c0 = mn, delta;
  if (nbins == 1)
    delta = 2 * (mx - mn);
  else
    delta = (mx - mn) / (double)(nbins - 1);
  double ib = 1.0 / delta;
  for (size_t i = 0; i < n; ++i) {
    uint32_t idx = (uint32_t)((d[i] - c0) * ib + 0.5);
    bins[idx] += 1;
  }
  *bincent0 = c0;
  *bindelta = delta;
  return ORC_OK;
}

/* ===================================================================== */
/* FFT with FFTW r2c/c2r semantics (calls at src/PDFactory.cc:1094,1199) */
/* ===================================================================== */

typedef struct orc_fft {
  unsigned n;     /* real length */
  unsigned m;     /* complex length n/2 */
  double *tw;     /* m/2 complex twiddles e^{-2 pi i k/m} */
  double *rtw;    /* m complex twiddles e^{-2 pi i k/n} */
  unsigned *rev;  /* bit reversal of [0, m) */
  double *work;   /* 2*m */
  /* extended-precision variant (precise mode) */
  long double *twl, *rtwl, *workl;
} orc_fft;

static void orc_fft_free(orc_fft *f) {
  if (!f) return;
  free(f->twl);
  free(f->rtwl);
  free(f->workl);
  free(f->tw);
  free(f->rtw);
  free(f->rev);
  free(f->work);
  free(f);
}

static orc_fft *orc_fft_new(unsigned n) {
  if (n < 4 || (n & (n - 1))) return NULL;
  orc_fft *f = (orc_fft *)calloc(1, sizeof(orc_fft));
  f->n = n;
  f->m = n / 2;
  unsigned m = f->m;
  f->tw = (double *)malloc(sizeof(double) * m);
  f->rtw = (double *)malloc(sizeof(double) * 2 * m);
  f->rev = (unsigned *)malloc(sizeof(unsigned) * m);
  f->work = (double *)malloc(sizeof(double) * 2 * m);
  /* twiddles from exact rational angles in extended precision */
  for (unsigned k = 0; k < m / 2; ++k) {
    long double a = -2.0L * (long double)ORC_PI * (long double)k / (long double)m;
    f->tw[2 * k] = (double)cosl(a);
    f->tw[2 * k + 1] = (double)sinl(a);
  }
  for (unsigned k = 0; k < m; ++k) {
    long double a = -2.0L * (long double)ORC_PI * (long double)k / (long double)n;
    f->rtw[2 * k] = (double)cosl(a);
    f->rtw[2 * k + 1] = (double)sinl(a);
  }
  unsigned lg = 0;
  while ((1u << lg) < m) ++lg;
  for (unsigned i = 0; i < m; ++i) {
    unsigned r = 0, v = i;
    for (unsigned b = 0; b < lg; ++b) {
      r = (r << 1) | (v & 1u);
      v >>= 1;
    }
    f->rev[i] = r;
  }
  return f;
}

/* in-place complex radix-2 DIT FFT of length m; sign -1 forward, +1
   backward (unnormalised) */
static void orc_cfft(const orc_fft *f, double *a, int sign) {
  unsigned m = f->m;
  for (unsigned i = 0; i < m; ++i) {
    unsigned r = f->rev[i];
    if (r > i) {
      double t0 = a[2 * i], t1 = a[2 * i + 1];
      a[2 * i] = a[2 * r];
      a[2 * i + 1] = a[2 * r + 1];
      a[2 * r] = t0;
      a[2 * r + 1] = t1;
    }
  }
  for (unsigned len = 2; len <= m; len <<= 1) {
    unsigned half = len >> 1, step = m / len;
    for (unsigned s = 0; s < m; s += len) {
      for (unsigned j = 0; j < half; ++j) {
        double wr = f->tw[2 * j * step];
        double wi = sign < 0 ? f->tw[2 * j * step + 1] : -f->tw[2 * j * step + 1];
        double *u = a + 2 * (s + j), *v = a + 2 * (s + j + half);
        double tr = v[0] * wr - v[1] * wi;
        double ti = v[0] * wi + v[1] * wr;
        v[0] = u[0] - tr;
        v[1] = u[1] - ti;
        u[0] = u[0] + tr;
        u[1] = u[1] + ti;
      }
    }
  }
}

static void orc_rfft_plan(const orc_fft *f, const double *in, double *out) {
  unsigned m = f->m;
  double *z = f->work;
  memcpy(z, in, sizeof(double) * 2 * m); /* z[j] = r[2j] + i r[2j+1] */
  orc_cfft(f, z, -1);
  for (unsigned k = 0; k <= m; ++k) {
    unsigned kk = (k == m) ? 0 : k, mk = (k == 0) ? 0 : m - k;
    double zr = z[2 * kk], zi = z[2 * kk + 1];
    double cr = z[2 * mk], ci = -z[2 * mk + 1]; /* conj(Z[m-k]) */
    double ar = 0.5 * (zr + cr), ai = 0.5 * (zi + ci);
    /* B = (Z - conj)/(2i) */
    double br = 0.5 * (zi - ci), bi = -0.5 * (zr - cr);
    double wr, wi;
    if (k == m) {
      wr = -1.0;
      wi = 0.0;
    } else {
      wr = f->rtw[2 * k];
      wi = f->rtw[2 * k + 1];
    }
    out[2 * k] = ar + (br * wr - bi * wi);
    out[2 * k + 1] = ai + (br * wi + bi * wr);
  }
}

static void orc_irfft_plan(const orc_fft *f, const double *in, double *out) {
  unsigned m = f->m;
  double *z = f->work;
  for (unsigned k = 0; k < m; ++k) {
    double xr = in[2 * k], xi = in[2 * k + 1];
    double yr = in[2 * (m - k)], yi = in[2 * (m - k) + 1];
    if (k == 0) { /* FFTW ignores Im of DC and Nyquist */
      xi = 0.0;
      yi = 0.0;
    }
    /* E = X + conj(Y);  O = (X - conj(Y)) * e^{+2 pi i k/n} */
    double er = xr + yr, ei = xi - yi;
    double dr = xr - yr, di = xi + yi;
    double wr = f->rtw[2 * k], wi = -f->rtw[2 * k + 1];
    double or_ = dr * wr - di * wi, oi = dr * wi + di * wr;
    /* Z = E + i O */
    z[2 * k] = er - oi;
    z[2 * k + 1] = ei + or_;
  }
  orc_cfft(f, z, +1);
  memcpy(out, z, sizeof(double) * 2 * m);
}

int orc_rfft(unsigned n, const double *in, double *out) {
  orc_fft *f = orc_fft_new(n);
  if (!f) return ORC_E_ARG;
  orc_rfft_plan(f, in, out);
  orc_fft_free(f);
  return ORC_OK;
}

int orc_irfft(unsigned n, const double *in, double *out) {
  orc_fft *f = orc_fft_new(n);
  if (!f) return ORC_E_ARG;
  orc_irfft_plan(f, in, out);
  orc_fft_free(f);
  return ORC_OK;
}

/* 2D real transforms of an n x n row-major array with FFTW's multi-dim
   r2c/c2r conventions (src/PDFactoryDouble.cc:727-781: the output is
   n x (n/2+1), the last dimension halved).  r2c: real transform of every
   row, then a complex transform down every column.  c2r: complex inverse
   down every column, then c2r along every row (imaginary parts of the
   k2 = 0 and n/2 entries of each row ignored, as FFTW's c2r does). */
int orc_rfft2(unsigned n, const double *in, double *out) {
  orc_fft *fr = orc_fft_new(n);     /* rows: real length n */
  orc_fft *fc = orc_fft_new(2 * n); /* columns: complex length n */
  if (!fr || !fc) {
    orc_fft_free(fr);
    orc_fft_free(fc);
    return ORC_E_ARG;
  }
  const unsigned nc = n / 2 + 1;
  for (unsigned i = 0; i < n; ++i) orc_rfft_plan(fr, in + (size_t)i * n, out + 2 * (size_t)i * nc);
  double *col = (double *)malloc(sizeof(double) * 2 * n);
  for (unsigned k2 = 0; k2 < nc; ++k2) {
    for (unsigned i = 0; i < n; ++i) {
      col[2 * i] = out[2 * ((size_t)i * nc + k2)];
      col[2 * i + 1] = out[2 * ((size_t)i * nc + k2) + 1];
    }
    orc_cfft(fc, col, -1);
    for (unsigned i = 0; i < n; ++i) {
      out[2 * ((size_t)i * nc + k2)] = col[2 * i];
      out[2 * ((size_t)i * nc + k2) + 1] = col[2 * i + 1];
    }
  }
  free(col);
  orc_fft_free(fr);
  orc_fft_free(fc);
  return ORC_OK;
}

int orc_irfft2(unsigned n, const double *in, double *out) {
  orc_fft *fr = orc_fft_new(n);
  orc_fft *fc = orc_fft_new(2 * n);
  if (!fr || !fc) {
    orc_fft_free(fr);
    orc_fft_free(fc);
    return ORC_E_ARG;
  }
  const unsigned nc = n / 2 + 1;
  double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)n * nc);
  double *col = (double *)malloc(sizeof(double) * 2 * n);
  for (unsigned k2 = 0; k2 < nc; ++k2) {
    for (unsigned i = 0; i < n; ++i) {
      col[2 * i] = in[2 * ((size_t)i * nc + k2)];
      col[2 * i + 1] = in[2 * ((size_t)i * nc + k2) + 1];
    }
    orc_cfft(fc, col, +1);
    for (unsigned i = 0; i < n; ++i) {
      tmp[2 * ((size_t)i * nc + k2)] = col[2 * i];
      tmp[2 * ((size_t)i * nc + k2) + 1] = col[2 * i + 1];
    }
  }
  for (unsigned i = 0; i < n; ++i) orc_irfft_plan(fr, tmp + 2 * (size_t)i * nc, out + (size_t)i * n);
  free(col);
  free(tmp);
  orc_fft_free(fr);
  orc_fft_free(fc);
  return ORC_OK;
}

/* Extended-precision (x87 long double, 64-bit significand) versions of the
   same transforms.  They are not the reference's arithmetic: they give the
   exact-math value of the reference's double-precision pipeline to ~1e-19
   of the peak, so that the intrinsic rounding noise of any double FFT (the
   reference's FFTW included) can be measured and excluded from parity. */
static void orc_fft_prepare_l(orc_fft *f) {
  if (f->twl) return;
  unsigned m = f->m, n = f->n;
  f->twl = (long double *)malloc(sizeof(long double) * m);
  f->rtwl = (long double *)malloc(sizeof(long double) * 2 * m);
  f->workl = (long double *)malloc(sizeof(long double) * 2 * m);
  for (unsigned k = 0; k < m / 2; ++k) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L *
                    (long double)k / (long double)m;
    f->twl[2 * k] = cosl(a);
    f->twl[2 * k + 1] = sinl(a);
  }
  for (unsigned k = 0; k < m; ++k) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L *
                    (long double)k / (long double)n;
    f->rtwl[2 * k] = cosl(a);
    f->rtwl[2 * k + 1] = sinl(a);
  }
}

static void orc_cfft_l(const orc_fft *f, long double *a, int sign) {
  unsigned m = f->m;
  for (unsigned i = 0; i < m; ++i) {
    unsigned r = f->rev[i];
    if (r > i) {
      long double t0 = a[2 * i], t1 = a[2 * i + 1];
      a[2 * i] = a[2 * r];
      a[2 * i + 1] = a[2 * r + 1];
      a[2 * r] = t0;
      a[2 * r + 1] = t1;
    }
  }
  for (unsigned len = 2; len <= m; len <<= 1) {
    unsigned half = len >> 1, step = m / len;
    for (unsigned s = 0; s < m; s += len) {
      for (unsigned j = 0; j < half; ++j) {
        long double wr = f->twl[2 * j * step];
        long double wi = sign < 0 ? f->twl[2 * j * step + 1] : -f->twl[2 * j * step + 1];
        long double *u = a + 2 * (s + j), *v = a + 2 * (s + j + half);
        long double tr = v[0] * wr - v[1] * wi;
        long double ti = v[0] * wi + v[1] * wr;
        v[0] = u[0] - tr;
        v[1] = u[1] - ti;
        u[0] = u[0] + tr;
        u[1] = u[1] + ti;
      }
    }
  }
}

static void orc_rfft_l(orc_fft *f, const double *in, long double *out) {
  orc_fft_prepare_l(f);
  unsigned m = f->m;
  long double *z = f->workl;
  for (unsigned i = 0; i < 2 * m; ++i) z[i] = in[i];
  orc_cfft_l(f, z, -1);
  for (unsigned k = 0; k <= m; ++k) {
    unsigned kk = (k == m) ? 0 : k, mk = (k == 0) ? 0 : m - k;
    long double zr = z[2 * kk], zi = z[2 * kk + 1];
    long double cr = z[2 * mk], ci = -z[2 * mk + 1];
    long double ar = 0.5L * (zr + cr), ai = 0.5L * (zi + ci);
    long double br = 0.5L * (zi - ci), bi = -0.5L * (zr - cr);
    long double wr, wi;
    if (k == m) {
      wr = -1.0L;
      wi = 0.0L;
    } else {
      wr = f->rtwl[2 * k];
      wi = f->rtwl[2 * k + 1];
    }
    out[2 * k] = ar + (br * wr - bi * wi);
    out[2 * k + 1] = ai + (br * wi + bi * wr);
  }
}

static void orc_irfft_l(orc_fft *f, const long double *in, double *out) {
  orc_fft_prepare_l(f);
  unsigned m = f->m;
  long double *z = f->workl;
  for (unsigned k = 0; k < m; ++k) {
    long double xr = in[2 * k], xi = in[2 * k + 1];
    long double yr = in[2 * (m - k)], yi = in[2 * (m - k) + 1];
    if (k == 0) {
      xi = 0.0L;
      yi = 0.0L;
    }
    long double er = xr + yr, ei = xi - yi;
    long double dr = xr - yr, di = xi + yi;
    long double wr = f->rtwl[2 * k], wi = -f->rtwl[2 * k + 1];
    long double or_ = dr * wr - di * wi, oi = dr * wi + di * wr;
    z[2 * k] = er - oi;
    z[2 * k + 1] = ei + or_;
  }
  orc_cfft_l(f, z, +1);
  for (unsigned i = 0; i < 2 * m; ++i) out[i] = (double)z[i];
}

/* ===================================================================== */
/* PD: src/PD.cc                                                          */
/* ===================================================================== */

double orc_pd_integral(unsigned n, const double *pd, double dflux,
                       int islog) {
  /* PD::getIntegral src/PD.cc:123-140 */
  if (n == 0) return NAN;
  double tot;
  if (islog) {
    tot = 0.5 * exp2(pd[0]);
    for (unsigned i = 1; i < n - 1; ++i) tot += exp2(pd[i]);
    tot += 0.5 * exp2(pd[n - 1]);
  } else {
    tot = 0.5 * pd[0];
    for (unsigned i = 1; i < n - 1; ++i) tot += pd[i];
    tot += 0.5 * pd[n - 1];
  }
  return tot * dflux;
}

double orc_pd_mean(unsigned n, const double *pd, double minflux, double dflux,
                   int islog, int donorm) {
  /* PD::getMean src/PD.cc:245-274 */
  if (n == 0) return NAN;
  double mean = 0.0;
  if (islog) {
    for (unsigned i = 1; i < n - 1; ++i) mean += (double)i * exp2(pd[i]);
    mean += 0.5 * (double)(n - 1) * exp2(pd[n - 1]);
  } else {
    for (unsigned i = 1; i < n - 1; ++i) mean += (double)i * pd[i];
    mean += 0.5 * (double)(n - 1) * pd[n - 1];
  }
  mean *= dflux * dflux;
  if (donorm) mean /= orc_pd_integral(n, pd, dflux, islog);
  mean += minflux;
  return mean;
}

void orc_pd_mean_var(unsigned n, const double *pd, double minflux,
                     double dflux, int islog, int donorm, double *pmean,
                     double *pvar) {
  /* PD::getMeanAndVar src/PD.cc:280-344 */
  if (n == 0) {
    *pmean = *pvar = NAN;
    return;
  }
  double normfac = 1.0;
  if (donorm) normfac = 1.0 / orc_pd_integral(n, pd, dflux, islog);
  double mean = 0.0;
  if (islog) {
    for (unsigned i = 1; i < n - 1; ++i) mean += (double)i * exp2(pd[i]);
    mean += 0.5 * (double)(n - 1) * exp2(pd[n - 1]);
  } else {
    for (unsigned i = 1; i < n - 1; ++i) mean += (double)i * pd[i];
    mean += 0.5 * (double)(n - 1) * pd[n - 1];
  }
  mean *= dflux * dflux;
  if (donorm) mean *= normfac;
  double var = 0.0, xval;
  if (islog) {
    for (unsigned i = 1; i < n - 1; ++i) {
      xval = (double)i * dflux - mean;
      var += xval * xval * exp2(pd[i]);
    }
    xval = (double)(n - 1) * dflux - mean;
    var += 0.5 * xval * xval * exp2(pd[n - 1]);
  } else {
    for (unsigned i = 1; i < n - 1; ++i) {
      xval = (double)i * dflux - mean;
      var += xval * xval * pd[i];
    }
    xval = (double)(n - 1) * dflux - mean;
    var += 0.5 * xval * xval * pd[n - 1];
  }
  var *= dflux;
  if (donorm) var *= normfac;
  mean += minflux;
  *pmean = mean;
  *pvar = var;
}

double orc_pd_interp(unsigned n, const double *pd, double minflux,
                     double dflux, double x) {
  /* PD::getPDVal(double) src/PD.cc:443-455 */
  int idx = (int)((x - minflux) / dflux);
  double maxfluxlm = minflux + (double)(n - 2) * dflux;
  if (x < minflux) return pd[0];
  if (x > maxfluxlm) return pd[n - 1];
  double p0 = pd[idx];
  double dx = x - (minflux + (double)idx * dflux);
  return p0 + dx / dflux * (pd[idx + 1] - p0);
}

double orc_pd_loglike_binned(unsigned n, const double *pd, double minflux,
                             double dflux, int islog, const uint32_t *bins,
                             uint32_t nbins, double bincent0,
                             double bindelta) {
  /* PD::getLogLikeBinned src/PD.cc:557-619 */
  double maxflux = minflux + (double)(n - 1) * dflux;
  double loglike = 0.0;
  double idflux = 1.0 / dflux;
  double maxfluxval = maxflux - dflux;
  for (uint32_t i = 0; i < nbins; ++i) {
    uint32_t ninbin = bins[i];
    if (ninbin == 0) continue;
    double cflux = bincent0 + (double)i * bindelta;
    if (islog) {
      if (cflux < minflux)
        loglike += (double)ninbin * pd[0];
      else if (cflux > maxfluxval)
        loglike += (double)ninbin * pd[n - 1];
      else {
        double delt = (cflux - minflux) * idflux;
        int idx = (int)delt;
        double t = delt - (double)idx;
        loglike += ((1.0 - t) * pd[idx] + t * pd[idx + 1]) * (double)ninbin;
      }
    } else {
      double delt = (cflux - minflux) * idflux;
      int idx = (int)delt;
      if (cflux < minflux)
        loglike += (double)ninbin * log2(pd[0]);
      else if (cflux > maxfluxval)
        loglike += (double)ninbin * log2(pd[n - 1]);
      else {
        double t = delt - (double)idx;
        loglike += ((1.0 - t) * log2(pd[idx]) + t * log2(pd[idx + 1])) *
                   (double)ninbin;
      }
    }
  }
  return orc_log2toe() * loglike;
}

double orc_pd_loglike_unbinned(unsigned n, const double *pd, double minflux,
                               double dflux, int islog, const double *flux,
                               size_t ndata, double *compensated) {
  /* PD::getLogLikeUnbinned src/PD.cc:628-677; the compensated (Neumaier)
     sum of the same terms is returned through *compensated. */
  double maxflux = minflux + (double)(n - 1) * dflux;
  double loglike = 0.0, s = 0.0, comp = 0.0;
  double idflux = 1.0 / dflux;
  double maxfluxval = maxflux - dflux;
  for (size_t i = 0; i < ndata; ++i) {
    double cflux = flux[i], term;
    if (islog) {
      if (cflux <= minflux)
        term = pd[0];
      else if (cflux >= maxfluxval)
        term = pd[n - 1];
      else {
        double delt = (cflux - minflux) * idflux;
        int idx = (int)delt;
        double t = delt - (double)idx;
        term = (1.0 - t) * pd[idx] + t * pd[idx + 1];
      }
    } else {
      if (cflux <= minflux)
        term = log2(pd[0]);
      else if (cflux >= maxfluxval)
        term = log2(pd[n - 1]);
      else {
        double delt = (cflux - minflux) * idflux;
        int idx = (int)delt;
        double t = delt - (double)idx;
        term = (1.0 - t) * log2(pd[idx]) + t * log2(pd[idx + 1]);
      }
    }
    loglike += term;
    double tt = s + term;
    if (fabs(s) >= fabs(term))
      comp += (s - tt) + term;
    else
      comp += (term - tt) + s;
    s = tt;
  }
  if (compensated) *compensated = orc_log2toe() * (s + comp);
  return orc_log2toe() * loglike;
}

/* ===================================================================== */
/* PDFactory (src/PDFactory.cc) working state                             */
/* ===================================================================== */

typedef struct orc_pdfac {
  unsigned currsize, ninterp;
  double mn, var_noi, sgpos, sgneg;
  double minflux_R, dflux;
  unsigned wrapRidx;
  int doshift;
  double shift;
  double r0;
  double *rvals, *prtrans, *pval, *pofd;
  double *RinterpFlux_pos, *RinterpVals_pos, *c_pos;
  double *RinterpFlux_neg, *RinterpVals_neg, *c_neg;
  int has_interp_pos, has_interp_neg;
  orc_fft *fft;
  int precise;
  long double *lprtrans, *lpval;
  /* last PD */
  double *pd;
  unsigned pd_n;
  double pd_minflux, pd_dflux;
  int pd_islog;
  unsigned last_split;
  double last_peak, last_integral;
  int split_shift; /* test hook: added to the split index found */
  /* the reference's TIMING categories (src/PDFactory.cc:93-125), thread
     CPU seconds, accumulated while `timing` is set */
  int timing;
  double tsec[ORC_NTIME];
} orc_pdfac;

static double orc_now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_THREAD_CPUTIME_ID, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* adds the time since *t0 to category k and restarts the clock */
static void orc_tick(orc_pdfac *p, double *t0, int k) {
  if (!p->timing) return;
  double t1 = orc_now();
  p->tsec[k] += t1 - *t0;
  *t0 = t1;
}
static double orc_tstart(const orc_pdfac *p) { return p->timing ? orc_now() : 0.0; }

static orc_pdfac *orc_pdfac_new(unsigned n, unsigned ninterp, int precise) {
  orc_pdfac *p = (orc_pdfac *)calloc(1, sizeof(orc_pdfac));
  p->precise = precise;
  if (precise) {
    p->lprtrans = (long double *)malloc(sizeof(long double) * 2 * (n / 2 + 1));
    p->lpval = (long double *)malloc(sizeof(long double) * 2 * (n / 2 + 1));
  }
  p->currsize = n;
  p->ninterp = ninterp;
  p->rvals = (double *)malloc(sizeof(double) * n);
  p->pofd = (double *)malloc(sizeof(double) * n);
  p->pd = (double *)malloc(sizeof(double) * n);
  p->prtrans = (double *)malloc(sizeof(double) * 2 * (n / 2 + 1));
  p->pval = (double *)malloc(sizeof(double) * 2 * (n / 2 + 1));
  p->RinterpFlux_pos = (double *)malloc(sizeof(double) * ninterp);
  p->RinterpVals_pos = (double *)malloc(sizeof(double) * ninterp);
  p->c_pos = (double *)malloc(sizeof(double) * ninterp);
  p->RinterpFlux_neg = (double *)malloc(sizeof(double) * ninterp);
  p->RinterpVals_neg = (double *)malloc(sizeof(double) * ninterp);
  p->c_neg = (double *)malloc(sizeof(double) * ninterp);
  p->fft = orc_fft_new(n);
  p->mn = p->var_noi = p->sgpos = p->sgneg = NAN;
  return p;
}

static void orc_pdfac_free(orc_pdfac *p) {
  if (!p) return;
  free(p->lprtrans);
  free(p->lpval);
  free(p->rvals);
  free(p->pofd);
  free(p->pd);
  free(p->prtrans);
  free(p->pval);
  free(p->RinterpFlux_pos);
  free(p->RinterpVals_pos);
  free(p->c_pos);
  free(p->RinterpFlux_neg);
  free(p->RinterpVals_neg);
  free(p->c_neg);
  orc_fft_free(p->fft);
  free(p);
}

/* PDFactory::initRInterp, src/PDFactory.cc:426-527 */
static int orc_init_rinterp(orc_pdfac *p, const orc_knots *model,
                            const orc_beam *bm, orc_err *err) {
  int has_pos = bm->haspos, has_neg = bm->hasneg;
  if (!(has_pos || has_neg))
    return orc_fail(err, ORC_E_ARG, "PDFactory", "computeR",
                    "Beam has neither positive nor negative bits");
  double modelmin = model->knots[0];
  double modelmax = model->knots[model->nknots - 1];
  unsigned ninterp = p->ninterp;
  double ininterpm1 = 1.0 / (double)(ninterp - 1);
  const double subedgemult = 1e-5; /* include/PDFactory.h:29 */
  p->has_interp_pos = p->has_interp_neg = 0;
  if (has_pos) {
    double mininterpflux = modelmin * subedgemult;
    double maxinterpflux = modelmax * bm->pospix[bm->npos - 1];
    double dinterp = 0.001 * (maxinterpflux - mininterpflux) * ininterpm1;
    maxinterpflux -= dinterp;
    double dinterpflux = (log2(maxinterpflux / mininterpflux)) * ininterpm1;
    for (unsigned i = 0; i < ninterp; ++i)
      p->RinterpFlux_pos[i] = mininterpflux * exp2((double)i * dinterpflux);
    orc_knots_getR(model, ninterp, p->RinterpFlux_pos, bm, p->RinterpVals_pos);
    for (unsigned i = 0; i < ninterp; ++i) {
      double val = p->RinterpVals_pos[i];
      if (val > 0)
        p->RinterpVals_pos[i] = log2(val);
      else
        p->RinterpVals_pos[i] = orc_smalllogval;
    }
    if (orc_cspline_init(ninterp, p->RinterpFlux_pos, p->RinterpVals_pos,
                         p->c_pos) != ORC_OK)
      return orc_fail(err, ORC_E_INTERNAL, "PDFactory", "initRInterp",
                      "gsl_spline_init failed (pos)");
    p->has_interp_pos = 1;
  }
  if (has_neg) {
    double maxinterpflux = -modelmin * subedgemult;
    double mininterpflux = -modelmax * bm->negpix[bm->nneg - 1];
    double dinterp = 0.001 * (maxinterpflux - mininterpflux) * ininterpm1;
    mininterpflux += dinterp;
    maxinterpflux -= dinterp;
    double dinterpflux = (log2(mininterpflux / maxinterpflux)) * ininterpm1;
    for (unsigned i = 0; i < ninterp; ++i)
      p->RinterpFlux_neg[i] =
          maxinterpflux * exp2((double)(ninterp - i - 1) * dinterpflux);
    orc_knots_getR(model, ninterp, p->RinterpFlux_neg, bm, p->RinterpVals_neg);
    for (unsigned i = 0; i < ninterp; ++i) {
      double val = p->RinterpVals_neg[i];
      if (val > 0)
        p->RinterpVals_neg[i] = log2(val);
      else
        p->RinterpVals_neg[i] = orc_smalllogval;
    }
    if (orc_cspline_init(ninterp, p->RinterpFlux_neg, p->RinterpVals_neg,
                         p->c_neg) != ORC_OK)
      return orc_fail(err, ORC_E_INTERNAL, "PDFactory", "initRInterp",
                      "gsl_spline_init failed (neg)");
    p->has_interp_neg = 1;
  }
  return ORC_OK;
}

/* PDFactory::initRFlux, src/PDFactory.cc:361-414 */
static int orc_init_rflux(orc_pdfac *p, unsigned n, double minflux,
                          double maxflux, orc_err *err) {
  if (n == 0)
    return orc_fail(err, ORC_E_ARG, "PDFactory", "initRFlux", "Invalid (0) n");
  if (maxflux < minflux) {
    double t = minflux;
    minflux = maxflux;
    maxflux = t;
  }
  if (n == 1) {
    if (minflux < 0.0)
      return orc_fail(err, ORC_E_ARG, "PDFactory", "initRFlux",
                      "n must be > 1 with negative fluxes");
    p->dflux = maxflux - minflux;
    p->minflux_R = minflux;
    p->wrapRidx = 0;
    return ORC_OK;
  }
  double inm1 = 1.0 / (double)(n - 1);
  if (maxflux > 0) {
    if (minflux > 0) minflux = 0.0;
    if (minflux == 0) {
      p->wrapRidx = n - 1;
      p->minflux_R = 0.0;
      p->dflux = maxflux * inm1;
    } else {
      double dflux = (maxflux - minflux) * inm1;
      dflux = maxflux / (n - floor(-minflux / dflux) - 1.0);
      p->dflux = dflux;
      p->wrapRidx = (unsigned)(maxflux / dflux + 0.9999999999);
      p->minflux_R = (p->wrapRidx + 1) * dflux - (double)n * dflux;
    }
  } else {
    if (minflux >= 0)
      return orc_fail(err, ORC_E_ARG, "PDFactory", "initRFlux",
                      "No effective range in R fluxes");
    p->dflux = -minflux * inm1;
    p->wrapRidx = 0;
    p->minflux_R = minflux;
  }
  return ORC_OK;
}

/* PDFactory::initR, src/PDFactory.cc:541-620 (muldr = true) */
static int orc_init_r(orc_pdfac *p, unsigned n, double minflux,
                      double maxflux, const orc_knots *model,
                      const orc_beam *bm, orc_err *err) {
  int rc = orc_init_rinterp(p, model, bm, err);
  if (rc) return rc;
  rc = orc_init_rflux(p, n, minflux, maxflux, err);
  if (rc) return rc;
  int haspos = bm->haspos && (maxflux > 0);
  int hasneg = bm->hasneg && (minflux < 0) && (p->wrapRidx > 0);
  unsigned ninterp = p->ninterp;
  double dflux = p->dflux;
  memset(p->rvals, 0, n * sizeof(double));
  if (haspos) {
    double minI = p->RinterpFlux_pos[0];
    double maxI = p->RinterpFlux_pos[ninterp - 1];
    unsigned minitidx = (unsigned)(minI / dflux + 0.9999999999999999);
    unsigned maxitidx = (unsigned)(maxI / dflux);
    if (maxitidx > p->wrapRidx) maxitidx = p->wrapRidx;
    for (unsigned i = minitidx; i <= maxitidx; ++i) {
      double cflux = (double)i * dflux;
      double splval = orc_cspline_eval(ninterp, p->RinterpFlux_pos,
                                       p->RinterpVals_pos, p->c_pos, cflux);
      p->rvals[i] = exp2(splval);
    }
  }
  if (hasneg) {
    double minI = p->RinterpFlux_neg[0];
    double maxI = p->RinterpFlux_neg[ninterp - 1];
    unsigned minitidx = (unsigned)((minI - p->minflux_R) / dflux +
                                   p->wrapRidx + 1 + 0.9999999999999999);
    if (minitidx <= p->wrapRidx) minitidx = p->wrapRidx + 1;
    unsigned maxitidx =
        (unsigned)((maxI - p->minflux_R) / dflux + p->wrapRidx + 1);
    if (maxitidx > n - 1) maxitidx = n - 1;
    double mval = p->minflux_R - (double)(p->wrapRidx + 1) * dflux;
    for (unsigned i = minitidx; i <= maxitidx; ++i) {
      double cflux = (double)i * dflux + mval;
      double splval = orc_cspline_eval(ninterp, p->RinterpFlux_neg,
                                       p->RinterpVals_neg, p->c_neg, cflux);
      p->rvals[i] = exp2(splval);
    }
  }
  for (unsigned i = 0; i < n; ++i) p->rvals[i] *= dflux;
  return ORC_OK;
}

/* PDFactory::getMeanVarFromR, src/PDFactory.cc:625-679 (rdflux = true) */
static void orc_mean_var_from_r(orc_pdfac *p) {
  double cf, prod;
  double mn = 0.0, var_noi = 0.0;
  unsigned wrapRidx = p->wrapRidx, currsize = p->currsize;
  double dflux = p->dflux;
  for (unsigned i = 1; i <= wrapRidx - 1; ++i) {
    cf = (double)i * dflux;
    prod = cf * p->rvals[i];
    mn += prod;
    var_noi += cf * prod;
  }
  cf = (double)wrapRidx * dflux;
  prod = 0.5 * cf * p->rvals[wrapRidx];
  mn += prod;
  var_noi += cf * prod;
  for (unsigned i = wrapRidx + 1; i < currsize; ++i) {
    cf = (double)(i - wrapRidx - 1) * dflux + p->minflux_R;
    prod = cf * p->rvals[i];
    mn += prod;
    var_noi += cf * prod;
  }
  if (wrapRidx < currsize - 1) {
    /* reference reads rvals[wrapRidx - 1] here (src/PDFactory.cc:664);
       reproduced as written */
    cf = p->minflux_R;
    prod = 0.5 * cf * p->rvals[wrapRidx - 1];
    mn -= prod;
    var_noi -= cf * prod;
  }
  p->mn = mn;
  p->var_noi = var_noi;
}

/* PDFactory::initPD, src/PDFactory.cc:1031-1141 */
static int orc_initpd(orc_pdfac *p, unsigned n, double minflux,
                      double maxflux, const orc_knots *model,
                      const orc_beam *bm, orc_err *err) {
  if (n == 0)
    return orc_fail(err, ORC_E_ARG, "PDFactory", "initPD", "Invalid (zero) n");
  if (n == 1)
    return orc_fail(err, ORC_E_ARG, "PDFactory", "initPD", "Invalid (==0) n");
  if (maxflux <= 0.0)
    return orc_fail(err, ORC_E_ARG, "PDFactory", "initPD",
                    "Invalid (non-positive) maxflux (%g)", maxflux);
  if (minflux == maxflux)
    return orc_fail(err, ORC_E_ARG, "PDFactory", "initPD",
                    "Invalid (0) range between min/max flux with value %g",
                    minflux);
  if (!orc_knots_is_valid(model))
    return orc_fail(err, ORC_E_INVALID_MODEL, "PDFactory", "initPD",
                    "Invalid model");
  if (!(bm->haspos || bm->hasneg))
    return orc_fail(err, ORC_E_ARG, "PDFactory", "initPD", "Empty beam");
  if (n != p->currsize || !p->fft)
    return orc_fail(err, ORC_E_ARG, "PDFactory", "setupTransforms",
                    "Transform size %u not supported by this oracle", n);
  double t0 = orc_tstart(p);
  int rc = orc_init_r(p, n, minflux, maxflux, model, bm, err);
  if (rc) return rc;
  orc_tick(p, &t0, ORC_T_R);
  orc_mean_var_from_r(p);
  orc_tick(p, &t0, ORC_T_OTHER);
  p->sgpos = sqrt(p->var_noi);
  p->doshift = (p->sgpos > p->dflux) && (fabs(p->mn) > p->dflux);
  p->shift = p->doshift ? -p->mn : 0.0;
  if (p->precise) {
    /* extended-precision transform and exponentiation; prtrans keeps the
       rounded values for inspection */
    long double *lp = p->lprtrans;
    orc_rfft_l(p->fft, p->rvals, lp);
    long double iflux_l = (long double)orc_two_pi / ((long double)n * p->dflux);
    long double r0l = lp[0];
    p->r0 = (double)r0l;
    if (p->precise == 3) {
      /* analysis mode: exact transform, but X_k and R^_0 rounded to double
         before the exponent difference is formed (as any double code does) */
      for (unsigned i = 0; i < 2 * (n / 2 + 1); ++i) lp[i] = (long double)(double)lp[i];
      r0l = (long double)(double)r0l;
    }
    lp[0] = 1.0L;
    lp[1] = 0.0L;
    long double wfac = (long double)p->shift * iflux_l;
    for (unsigned idx = 1; idx < n / 2 + 1; ++idx) {
      long double expfac = expl(lp[2 * idx] - r0l);
      long double ival = lp[2 * idx + 1] - (p->doshift ? wfac * (long double)idx : 0.0L);
      lp[2 * idx] = expfac * cosl(ival);
      lp[2 * idx + 1] = expfac * sinl(ival);
    }
    for (unsigned i = 0; i < 2 * (n / 2 + 1); ++i) p->prtrans[i] = (double)lp[i];
    return ORC_OK;
  }
  orc_rfft_plan(p->fft, p->rvals, p->prtrans);
  orc_tick(p, &t0, ORC_T_FFT);
  double *pr = p->prtrans;
  double iflux = orc_two_pi / (n * p->dflux);
  unsigned ncplx = n / 2 + 1;
  double r0 = pr[0];
  p->r0 = r0;
  pr[0] = 1.0;
  pr[1] = 0.0;
  if (p->doshift) {
    double wfac = p->shift * iflux;
    for (unsigned idx = 1; idx < ncplx; ++idx) {
      double expfac = exp(pr[2 * idx] - r0);
      double ival = pr[2 * idx + 1] - wfac * (double)idx;
      pr[2 * idx] = expfac * cos(ival);
      pr[2 * idx + 1] = expfac * sin(ival);
    }
  } else {
    for (unsigned idx = 1; idx < ncplx; ++idx) {
      double expfac = exp(pr[2 * idx] - r0);
      double ival = pr[2 * idx + 1];
      pr[2 * idx] = expfac * cos(ival);
      pr[2 * idx + 1] = expfac * sin(ival);
    }
  }
  orc_tick(p, &t0, ORC_T_P0);
  return ORC_OK;
}

/* PDFactory::findSplitPoint, src/PDFactory.cc:691-921 */
static int orc_find_split(const orc_pdfac *p, unsigned *out, double *peakout,
                          orc_err *err) {
  const double n1 = 2.5, f1 = 5e-14, f2 = 1e-5, f3 = 5;
  const double *pofd = p->pofd;
  unsigned currsize = p->currsize;
  double dflux = p->dflux, sgpos = p->sgpos, sgneg = p->sgneg;
  if (sgpos <= 0)
    return orc_fail(err, ORC_E_SPLIT, "PDFactory", "findSplitPoint",
                    "Invalid (non-positive) sigma positive");
  if (sgneg <= 0)
    return orc_fail(err, ORC_E_SPLIT, "PDFactory", "findSplitPoint",
                    "Invalid (non-positive) sigma negative");
  double fluxrange = currsize * dflux;
  double n1sigpos = n1 * sgpos;
  double n1signeg = n1 * sgneg;
  if (fluxrange < (n1sigpos + n1signeg))
    return orc_fail(err, ORC_E_SPLIT, "PDFactory", "findSplitPoint",
                    "Not enough flux range; range is: %g sigmas are: %g %g",
                    fluxrange, sgpos, sgneg);
  double peak = pofd[0], currval;
  for (unsigned i = 1; i < currsize; ++i) {
    currval = pofd[i];
    if (currval > peak) peak = currval;
  }
  *peakout = peak;
  if (peak <= 0.0)
    return orc_fail(err, ORC_E_SPLIT, "PDFactory", "findSplitPoint",
                    "Peak P(D) value was zero");
  int botidx, topidx;
  botidx = (int)(n1sigpos / dflux);
  if (n1sigpos > fabs(p->minflux_R))
    topidx = (int)p->wrapRidx;
  else
    topidx = (int)(-(n1signeg + p->minflux_R) / dflux + p->wrapRidx + 1);
  if (botidx >= (int)currsize)
    return orc_fail(err, ORC_E_SPLIT, "PDFactory", "findSplitPoint",
                    "Logic error; invalid botidx (wrapped) in split search");
  if (topidx >= (int)currsize)
    return orc_fail(err, ORC_E_SPLIT, "PDFactory", "findSplitPoint",
                    "Logic error; invalid (large) topidx in split search");
  int minidx = botidx;
  double minval = pofd[minidx];
  for (int i = botidx + 1; i <= topidx; ++i) {
    currval = pofd[i];
    if (currval < minval) {
      minidx = i;
      minval = currval;
    }
  }
  double targval = f2 * peak;
  if (minval > targval)
    return orc_fail(err, ORC_E_SPLIT, "PDFactory", "findSplitPoint",
                    "Minimum in pd is too large relative to peak; can't find "
                    "split point with minimum: %g peak value: %g Required "
                    "split value < %g in index range %d to %d in P(D) of "
                    "size: %u",
                    minval, peak, targval, botidx, topidx, currsize);
  const int pre_jitter_n = 17;
  double jitter = 0.0;
  if (minidx > pre_jitter_n) {
    int monotonic_decrease = 1;
    int initidx = minidx - pre_jitter_n;
    double prevval;
    jitter = prevval = pofd[initidx];
    for (int i = initidx + 1; i < minidx; ++i) {
      currval = pofd[i];
      if (currval >= prevval) monotonic_decrease = 0;
      jitter += currval;
      prevval = currval;
    }
    if (monotonic_decrease) {
      *out = (unsigned)minidx;
      return ORC_OK;
    } else {
      jitter += minval;
      jitter = 1.1 * jitter / (double)(pre_jitter_n + 1);
    }
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
    if (minidx < botidx) bi = botidx; else bi = minidx - nminup;
    if (minidx > topidx - nminup) ti = topidx - nminup; else ti = minidx + nminup;
    for (int i = bi; i < ti; ++i) {
      currval = pofd[i];
      if (currval > minval) minval = currval;
    }
  }
  if (minval > 0) jitter = (f3 * minval) < jitter ? (f3 * minval) : jitter;
  if ((botidx >= topidx) || (abs(topidx - botidx) < 1))
    return orc_fail(err, ORC_E_SPLIT, "PDFactory", "findSplitPoint",
                    "Topidx (%d) vs. botidx (%d) search issue -- can't find "
                    "range to search",
                    topidx, botidx);
  if (pofd[topidx] < jitter)
    return orc_fail(err, ORC_E_SPLIT, "PDFactory", "findSplitPoint",
                    "Topidx (%d) is already below jitter estimate (%g) with "
                    "value %g minimum: %g peak: %g; can't find split",
                    topidx, jitter, pofd[topidx], minval, peak);
  for (int i = topidx - 1; i >= botidx; --i)
    if (pofd[i] < jitter) {
      *out = (unsigned)(i + 1);
      return ORC_OK;
    }
  minval = pofd[topidx];
  for (int i = topidx - 1; i >= botidx; --i)
    if (pofd[i] < minval) minval = pofd[i];
  return orc_fail(err, ORC_E_SPLIT, "PDFactory", "findSplitPoint",
                  "Unable to find split point in range %d down to %d with "
                  "value less than %g; minimum value encountered: %g",
                  topidx, botidx, jitter, minval);
}

/* findSplitPoint on a caller's array (src/PDFactory.cc:691-921), for
   hand-built cases: the members the reference reads (currsize, dflux,
   sgpos, sgneg, minflux_R, wrapRidx) are given directly */
int orc_find_split_array(const double *pofd, unsigned n, double dflux,
                         double sgpos, double sgneg, double minflux_R,
                         unsigned wrapRidx, unsigned *out, double *peak,
                         orc_err *err) {
  orc_pdfac p;
  memset(&p, 0, sizeof p);
  p.pofd = (double *)pofd;
  p.currsize = n;
  p.dflux = dflux;
  p.sgpos = sgpos;
  p.sgneg = sgneg;
  p.minflux_R = minflux_R;
  p.wrapRidx = wrapRidx;
  return orc_find_split(&p, out, peak, err);
}

/* PDFactory::unwrapAndNormalizePD, src/PDFactory.cc:934-1009 */
static int orc_unwrap_normalize(orc_pdfac *p, orc_err *err) {
  unsigned n = p->currsize;
  double t0 = orc_tstart(p);
  for (unsigned i = 0; i < n; ++i)
    if (p->pofd[i] < 0) p->pofd[i] = 0.0;
  orc_tick(p, &t0, ORC_T_POS);
  unsigned splitidx = 0;
  double peak = 0.0;
  int rc = orc_find_split(p, &splitidx, &peak, err);
  p->last_peak = peak;
  if (rc) return rc;
  orc_tick(p, &t0, ORC_T_SPLIT);
  if (p->split_shift) {
    /* test hook (logL insensitivity to the split inside the jitter floor) */
    long sh = (long)splitidx + p->split_shift;
    if (sh < 1 || sh >= (long)n)
      return orc_fail(err, ORC_E_ARG, "PDFactory", "unwrapAndNormalizePD",
                      "split shift leaves the array");
    splitidx = (unsigned)sh;
  }
  p->last_split = splitidx;
  memcpy(p->pd, p->pofd + splitidx, (n - splitidx) * sizeof(double));
  memcpy(p->pd + n - splitidx, p->pofd, splitidx * sizeof(double));
  p->pd_n = n;
  p->pd_islog = 0;
  p->pd_minflux = 0.0;
  p->pd_dflux = p->dflux;
  orc_tick(p, &t0, ORC_T_COPY);
  /* PD::normalize src/PD.cc:146-161 */
  double tot, tmn;
  if (p->precise) {
    /* the exact-math value: the trapezoid sums and the mean accumulated in
       extended precision, so no tail term is lost against the running sum
       (the reference's sequential double sums drop terms below half an ulp
       of the total) */
    const double *q = p->pd;
    long double lt = 0.5L * (long double)q[0];
    for (unsigned i = 1; i < n - 1; ++i) lt += (long double)q[i];
    lt += 0.5L * (long double)q[n - 1];
    lt *= (long double)p->pd_dflux;
    tot = (double)lt;
    p->last_integral = tot;
    long double ilt = 1.0L / lt, lm = 0.0L;
    for (unsigned i = 1; i < n - 1; ++i) lm += (long double)i * ((long double)q[i] * ilt);
    lm += 0.5L * (long double)(n - 1) * ((long double)q[n - 1] * ilt);
    lm *= (long double)p->pd_dflux * (long double)p->pd_dflux;
    for (unsigned i = 0; i < n; ++i) p->pd[i] = (double)((long double)q[i] * ilt);
    tmn = (double)lm;
  } else {
    tot = orc_pd_integral(n, p->pd, p->pd_dflux, 0);
    p->last_integral = tot;
    double itot = 1.0 / tot;
    for (unsigned i = 0; i < n; ++i) p->pd[i] *= itot;
    orc_tick(p, &t0, ORC_T_NORM);
    tmn = orc_pd_mean(n, p->pd, 0.0, p->pd_dflux, 0, 0);
  }
  orc_tick(p, &t0, ORC_T_MEAN);
  if (isinf(tmn) || isnan(tmn))
    return orc_fail(err, ORC_E_MEAN, "PDFactory", "unwrapandNormalizePD",
                    "Un-shift amounts not finite: %g At length: %u with noise "
                    "estimates: %g %g",
                    tmn, n, p->sgpos, p->sgneg);
  p->pd_minflux = -tmn;
  return ORC_OK;
}

/* PDFactory::getPD, src/PDFactory.cc:1154-1243 */
static int orc_getpd(orc_pdfac *p, double sigma, int setlog, orc_err *err) {
  unsigned n = p->currsize;
  unsigned ncplx = n / 2 + 1;
  if (p->precise == 2) {
    /* analysis mode: extended forward pipeline, double inverse transform */
    double iflux = orc_two_pi / (n * p->dflux);
    double sigfac = -0.5 * sigma * sigma;
    for (unsigned idx = 0; idx < ncplx; ++idx) {
      double w = iflux * (double)idx;
      double ef = (sigma > 0 && idx > 0) ? exp(sigfac * w * w) : 1.0;
      p->pval[2 * idx] = ef * p->prtrans[2 * idx];
      p->pval[2 * idx + 1] = ef * p->prtrans[2 * idx + 1];
    }
    orc_irfft_plan(p->fft, p->pval, p->pofd);
  } else if (p->precise) {
    long double *lv = p->lpval, *lp = p->lprtrans;
    long double iflux = (long double)orc_two_pi / ((long double)n * p->dflux);
    long double sigfac = -0.5L * (long double)sigma * (long double)sigma;
    for (unsigned idx = 0; idx < ncplx; ++idx) {
      long double w = iflux * (long double)idx;
      long double ef = (sigma > 0 && idx > 0) ? expl(sigfac * w * w) : 1.0L;
      lv[2 * idx] = ef * lp[2 * idx];
      lv[2 * idx + 1] = ef * lp[2 * idx + 1];
      if (p->precise == 4) { /* analysis mode: spectrum rounded to double */
        lv[2 * idx] = (long double)(double)lv[2 * idx];
        lv[2 * idx + 1] = (long double)(double)lv[2 * idx + 1];
      }
    }
    orc_irfft_l(p->fft, lv, p->pofd);
  } else if (sigma > 0) {
    double t0 = orc_tstart(p);
    double iflux = orc_two_pi / (n * p->dflux);
    double sigfac = -0.5 * sigma * sigma;
    p->pval[0] = 1.0;
    p->pval[1] = 0.0;
    for (unsigned idx = 1; idx < ncplx; ++idx) {
      double w = iflux * (double)idx;
      double expfac = exp(sigfac * w * w);
      p->pval[2 * idx] = expfac * p->prtrans[2 * idx];
      p->pval[2 * idx + 1] = expfac * p->prtrans[2 * idx + 1];
    }
    orc_tick(p, &t0, ORC_T_P0);
    orc_irfft_plan(p->fft, p->pval, p->pofd);
    orc_tick(p, &t0, ORC_T_FFT);
  } else {
    double t0 = orc_tstart(p);
    orc_irfft_plan(p->fft, p->prtrans, p->pofd);
    orc_tick(p, &t0, ORC_T_FFT);
  }
  p->sgpos = sqrt(p->var_noi + sigma * sigma);
  if (9 * sigma * sigma > p->var_noi)
    p->sgneg = sigma;
  else
    p->sgneg = p->sgpos;
  int rc = orc_unwrap_normalize(p, err);
  if (rc) return rc;
  if (setlog) {
    double t0 = orc_tstart(p);
    /* PD::applyLog(false) src/PD.cc:169-184 */
    for (unsigned i = 0; i < n; ++i) {
      double val = p->pd[i];
      p->pd[i] = (val > 0.0) ? log2(val) : orc_smalllogval;
    }
    p->pd_islog = 1;
    orc_tick(p, &t0, ORC_T_LOG);
  }
  return ORC_OK;
}

/* ===================================================================== */
/* calcLikeSingle / calcLike: src/calcLike.cc                             */
/* ===================================================================== */

typedef struct orc_field {
  double *pix;
  size_t npix;
  uint32_t *bins;
  uint32_t nbins;
  double bincent0, bindelta;
  double sigma_base, like_norm, like_offset;
  double mindata, maxdata;
} orc_field;

typedef struct orc_beamset {
  orc_beam bm;
  orc_field *fields;
  unsigned nfields;
  double minRFlux, maxRFlux;
  double maxsigma_base, exp_conf;
  double minDataFlux, maxDataFlux;
} orc_beamset;

struct orc_like {
  unsigned fftsize, ninterp;
  orc_knots model;
  orc_beamset *sets;
  unsigned nsets;
  orc_err err;
  orc_pdfac *fac; /* single-threaded path workspace */
  int setup_done;
  int precise;
  int split_shift; /* test hook, see orc_pdfac */
  int timing;      /* per-stage timing of the batch workers */
  double tsec[ORC_NTIME];
  pthread_mutex_t tlock;
};

orc_like *orc_like_new(unsigned fftsize, unsigned ninterp) {
  orc_like *L = (orc_like *)calloc(1, sizeof(orc_like));
  pthread_mutex_init(&L->tlock, NULL);
  L->fftsize = fftsize;
  L->ninterp = ninterp;
  return L;
}

void orc_like_free(orc_like *L) {
  if (!L) return;
  for (unsigned b = 0; b < L->nsets; ++b) {
    orc_beamset *s = &L->sets[b];
    free(s->bm.pospix);
    free(s->bm.negpix);
    free(s->bm.posinv);
    free(s->bm.neginv);
    free(s->bm.posweights);
    free(s->bm.poshistval);
    free(s->bm.negweights);
    free(s->bm.neghistval);
    for (unsigned f = 0; f < s->nfields; ++f) {
      free(s->fields[f].pix);
      free(s->fields[f].bins);
    }
    free(s->fields);
  }
  free(L->sets);
  orc_pdfac_free(L->fac);
  free(L);
}

const orc_err *orc_like_error(const orc_like *L) { return &L->err; }

void orc_like_set_precise(orc_like *L, int precise) {
  L->precise = precise;
  orc_pdfac_free(L->fac);
  L->fac = NULL;
}

int orc_like_set_knot_positions(orc_like *L, unsigned n, const double *pos) {
  return orc_knots_set_positions(&L->model, n, pos, &L->err);
}

int orc_like_add_beamset(orc_like *L, const orc_beam *bm) {
  L->sets = (orc_beamset *)realloc(L->sets, sizeof(orc_beamset) * (L->nsets + 1));
  orc_beamset *s = &L->sets[L->nsets];
  memset(s, 0, sizeof(*s));
  orc_beam_copy(&s->bm, bm);
  s->minRFlux = s->maxRFlux = NAN;
  return (int)(L->nsets++);
}

int orc_like_add_field(orc_like *L, int b, const double *pix, size_t npix,
                       const uint32_t *bins, uint32_t nbins, double bincent0,
                       double bindelta, double sigma_base, double like_norm) {
  if (b < 0 || (unsigned)b >= L->nsets) return ORC_E_ARG;
  orc_beamset *s = &L->sets[b];
  s->fields = (orc_field *)realloc(s->fields, sizeof(orc_field) * (s->nfields + 1));
  orc_field *f = &s->fields[s->nfields];
  memset(f, 0, sizeof(*f));
  f->npix = npix;
  if (pix && npix) {
    f->pix = (double *)malloc(sizeof(double) * npix);
    memcpy(f->pix, pix, sizeof(double) * npix);
  }
  if (bins) {
    f->nbins = nbins;
    f->bins = (uint32_t *)malloc(sizeof(uint32_t) * nbins);
    memcpy(f->bins, bins, sizeof(uint32_t) * nbins);
    f->bincent0 = bincent0;
    f->bindelta = bindelta;
    /* fitsData::getMinMax binned branch src/fitsData.cc:353-357 */
    f->mindata = bincent0;
    f->maxdata = bincent0 + (double)(nbins - 1) * bindelta;
  } else {
    orc_data_minmax(f->pix, npix, &f->mindata, &f->maxdata);
  }
  f->sigma_base = sigma_base;
  f->like_norm = like_norm;
  s->nfields++;
  return (int)(s->nfields - 1);
}

int orc_like_finalize_setup(orc_like *L, const float *theta, unsigned npar,
                            double exp_conf) {
  for (unsigned b = 0; b < L->nsets; ++b) {
    orc_beamset *s = &L->sets[b];
    if (s->nfields == 0)
      return orc_fail(&L->err, ORC_E_ARG, "calcLikeSingle",
                      "readDataFromFile", "No data sets");
    /* data range src/calcLike.cc:205-230 */
    s->minDataFlux = s->fields[0].mindata;
    s->maxDataFlux = s->fields[0].maxdata;
    for (unsigned f = 1; f < s->nfields; ++f) {
      if (s->fields[f].mindata < s->minDataFlux) s->minDataFlux = s->fields[f].mindata;
      if (s->fields[f].maxdata > s->maxDataFlux) s->maxDataFlux = s->fields[f].maxdata;
    }
    /* setExpConf + setSigmaBase (vector form) src/calcLike.cc:283-308 */
    s->exp_conf = exp_conf;
    s->maxsigma_base = s->fields[0].sigma_base;
    for (unsigned f = 1; f < s->nfields; ++f)
      if (s->fields[f].sigma_base > s->maxsigma_base)
        s->maxsigma_base = s->fields[f].sigma_base;
    for (unsigned f = 0; f < s->nfields; ++f) {
      double var = s->fields[f].sigma_base * s->fields[f].sigma_base +
                   exp_conf * exp_conf;
      s->fields[f].like_offset = 0.0;
      if (var > 0) {
        double ndd = (double)s->fields[f].npix;
        s->fields[f].like_offset =
            -0.5 * ndd * (log(orc_two_pi * var) + 1.0);
      }
    }
  }
  /* calcLike::setRRanges src/calcLike.cc:827-831 ->
     calcLikeSingle::setRRange src/calcLike.cc:346-371 */
  int rc = orc_knots_set_params(&L->model, npar, theta, &L->err);
  if (rc) return rc;
  if (!orc_knots_is_valid(&L->model))
    return orc_fail(&L->err, ORC_E_INVALID_MODEL, "numberCountsKnots",
                    "getRRange", "Model is not valid");
  for (unsigned b = 0; b < L->nsets; ++b) {
    orc_beamset *s = &L->sets[b];
    const double safetyfac = 1.05;
    double rmin, rmax;
    orc_knots_rrange(&L->model, &s->bm, &rmin, &rmax);
    double sigma = sqrt(s->exp_conf * s->exp_conf +
                        s->maxsigma_base * s->maxsigma_base);
    s->minRFlux = safetyfac * rmin;
    s->maxRFlux = safetyfac * rmax + orc_n_zero_pad * sigma;
    if (s->minDataFlux < s->minRFlux && s->bm.hasneg) s->minRFlux = s->minDataFlux;
    if (s->maxDataFlux > s->maxRFlux && s->bm.haspos) s->maxRFlux = s->maxDataFlux;
  }
  L->setup_done = 1;
  return ORC_OK;
}

int orc_like_beamset_info(const orc_like *L, int b, double *out) {
  if (b < 0 || (unsigned)b >= L->nsets) return ORC_E_ARG;
  const orc_beamset *s = &L->sets[b];
  out[0] = s->minRFlux;
  out[1] = s->maxRFlux;
  out[2] = s->maxsigma_base;
  out[3] = s->minDataFlux;
  out[4] = s->maxDataFlux;
  out[5] = s->nfields;
  out[6] = s->exp_conf;
  out[7] = 0;
  return ORC_OK;
}

static double orc_field_loglike_pd(const orc_pdfac *p, const orc_field *f,
                                   double *compensated) {
  if (f->bins) {
    double v = orc_pd_loglike_binned(p->pd_n, p->pd, p->pd_minflux,
                                     p->pd_dflux, p->pd_islog, f->bins,
                                     f->nbins, f->bincent0, f->bindelta);
    if (compensated) *compensated = v;
    return v;
  }
  return orc_pd_loglike_unbinned(p->pd_n, p->pd, p->pd_minflux, p->pd_dflux,
                                 p->pd_islog, f->pix, f->npix, compensated);
}

/* calcLikeSingle::getLogLike src/calcLike.cc:389-443 + calcLike::getLogLike
   src/calcLike.cc:884-955 (priors and bonus integrals are not enabled in
   this restatement: the measured configurations carry none). */
static int orc_loglike_ws(const orc_like *L, orc_knots *model, orc_pdfac *p,
                          const float *theta, unsigned npar, double *out,
                          orc_err *err) {
  if (L->nsets == 0) {
    *out = NAN;
    return ORC_OK;
  }
  if (npar < 2)
    return orc_fail(err, ORC_E_ARG, "calcLike", "getLogLike",
                    "Not enough elements in params");
  unsigned nknots = model->nknots;
  if (nknots == 0)
    return orc_fail(err, ORC_E_ARG, "calcLike", "getLogLike",
                    "Model knot positions not loaded");
  if (nknots > npar - 1)
    return orc_fail(err, ORC_E_ARG, "calcLike", "getLogLike",
                    "Not enough elements in paramSet");
  int rc = orc_knots_set_params(model, npar, theta, err);
  if (rc) return rc;
  double sigmult = (double)theta[nknots];
  double LogLike = 0.0;
  for (unsigned b = 0; b < L->nsets; ++b) {
    const orc_beamset *s = &L->sets[b];
    double max_sigma = s->maxsigma_base * sigmult;
    if (max_sigma < 0.0) {
      /* calcLikeSingle::bad_like (include/calcLike.h:71) is returned as
         that beam set's log likelihood */
      LogLike += 1e25;
      continue;
    }
    rc = orc_initpd(p, L->fftsize, s->minRFlux, s->maxRFlux, model, &s->bm, err);
    if (rc) return rc;
    for (unsigned f = 0; f < s->nfields; ++f) {
      rc = orc_getpd(p, sigmult * s->fields[f].sigma_base, 1, err);
      if (rc) return rc;
      double t0 = orc_tstart(p);
      double comp = 0.0;
      double curr = orc_field_loglike_pd(p, &s->fields[f], &comp);
      if (p->precise) curr = comp;  /* compensated pixel sum */
      orc_tick(p, &t0, ORC_T_LIKE);
      LogLike += (curr - s->fields[f].like_offset) / s->fields[f].like_norm;
    }
  }
  *out = LogLike;
  return ORC_OK;
}

static orc_pdfac *orc_like_fac(orc_like *L) {
  if (!L->fac) L->fac = orc_pdfac_new(L->fftsize, L->ninterp, L->precise);
  L->fac->split_shift = L->split_shift;
  return L->fac;
}

void orc_like_set_split_shift(orc_like *L, int shift) { L->split_shift = shift; }

int orc_like_loglike(orc_like *L, const float *theta, unsigned npar,
                     double *loglike) {
  memset(&L->err, 0, sizeof L->err);
  if (!L->setup_done)
    return orc_fail(&L->err, ORC_E_ARG, "calcLikeSingle", "getNegLogLike",
                    "R range not set");
  return orc_loglike_ws(L, &L->model, orc_like_fac(L), theta, npar, loglike,
                        &L->err);
}

typedef struct orc_job {
  orc_like *L;
  const float *theta;
  unsigned nwalk, stride, npar, tid, nthreads;
  double *loglike;
  int32_t *status;
} orc_job;

static void *orc_worker(void *arg) {
  orc_job *j = (orc_job *)arg;
  orc_pdfac *p = orc_pdfac_new(j->L->fftsize, j->L->ninterp, j->L->precise);
  p->split_shift = j->L->split_shift;
  p->timing = j->L->timing;
  orc_knots model = j->L->model;
  orc_err err;
  for (unsigned w = j->tid; w < j->nwalk; w += j->nthreads) {
    memset(&err, 0, sizeof err);
    double v = NAN;
    int rc = orc_loglike_ws(j->L, &model, p, j->theta + (size_t)w * j->stride,
                            j->npar, &v, &err);
    j->loglike[w] = rc ? NAN : v;
    j->status[w] = rc;
  }
  if (p->timing) {
    pthread_mutex_lock(&j->L->tlock);
    for (int k = 0; k < ORC_NTIME; ++k) j->L->tsec[k] += p->tsec[k];
    pthread_mutex_unlock(&j->L->tlock);
  }
  orc_pdfac_free(p);
  return NULL;
}

void orc_like_set_timing(orc_like *L, int on) {
  L->timing = on;
  memset(L->tsec, 0, sizeof L->tsec);
}

void orc_like_timing(const orc_like *L, double *out) {
  memcpy(out, L->tsec, sizeof L->tsec);
}

int orc_like_loglike_batch(orc_like *L, const float *theta, unsigned nwalk,
                           unsigned stride, unsigned npar, double *loglike,
                           int32_t *status, unsigned nthreads) {
  if (!L->setup_done) return ORC_E_ARG;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  orc_job jobs[256];
  for (unsigned t = 0; t < nthreads; ++t) {
    jobs[t] = (orc_job){L, theta, nwalk, stride, npar, t, nthreads, loglike, status};
    if (nthreads == 1)
      orc_worker(&jobs[t]);
    else
      pthread_create(&th[t], NULL, orc_worker, &jobs[t]);
  }
  if (nthreads > 1)
    for (unsigned t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  return ORC_OK;
}

int orc_like_initpd(orc_like *L, int b, const float *theta, unsigned npar,
                    double *rinterp_flux, double *rinterp_vals, double *rvals,
                    double *prtrans, double *scalars) {
  memset(&L->err, 0, sizeof L->err);
  if (b < 0 || (unsigned)b >= L->nsets || !L->setup_done) return ORC_E_ARG;
  int rc = orc_knots_set_params(&L->model, npar, theta, &L->err);
  if (rc) return rc;
  orc_pdfac *p = orc_like_fac(L);
  const orc_beamset *s = &L->sets[b];
  rc = orc_initpd(p, L->fftsize, s->minRFlux, s->maxRFlux, &L->model, &s->bm,
                  &L->err);
  if (rc) return rc;
  unsigned n = L->fftsize;
  if (rinterp_flux) memcpy(rinterp_flux, p->RinterpFlux_pos, sizeof(double) * L->ninterp);
  if (rinterp_vals) memcpy(rinterp_vals, p->RinterpVals_pos, sizeof(double) * L->ninterp);
  if (rvals) memcpy(rvals, p->rvals, sizeof(double) * n);
  if (prtrans) memcpy(prtrans, p->prtrans, sizeof(double) * 2 * (n / 2 + 1));
  if (scalars) {
    scalars[0] = p->dflux;
    scalars[1] = p->minflux_R;
    scalars[2] = p->wrapRidx;
    scalars[3] = p->mn;
    scalars[4] = p->var_noi;
    scalars[5] = p->doshift;
    scalars[6] = p->shift;
    scalars[7] = p->r0;
  }
  return ORC_OK;
}

int orc_like_rinterp_neg(orc_like *L, double *flux, double *vals) {
  /* negative-beam interpolation table of the last orc_like_initpd
     (RinterpFlux_neg / RinterpVals_neg, src/PDFactory.cc:486-524) */
  orc_pdfac *p = orc_like_fac(L);
  if (!p->has_interp_neg) return ORC_E_ARG;
  if (flux) memcpy(flux, p->RinterpFlux_neg, sizeof(double) * L->ninterp);
  if (vals) memcpy(vals, p->RinterpVals_neg, sizeof(double) * L->ninterp);
  return ORC_OK;
}

int orc_like_getpd(orc_like *L, int b, int f, double sigmult, int setlog,
                   double *pofd_raw, double *pd_out, double *info) {
  memset(&L->err, 0, sizeof L->err);
  if (b < 0 || (unsigned)b >= L->nsets || !L->fac) return ORC_E_ARG;
  const orc_beamset *s = &L->sets[b];
  if (f < 0 || (unsigned)f >= s->nfields) return ORC_E_ARG;
  orc_pdfac *p = L->fac;
  int rc = orc_getpd(p, sigmult * s->fields[f].sigma_base, setlog, &L->err);
  unsigned n = L->fftsize;
  if (pofd_raw) memcpy(pofd_raw, p->pofd, sizeof(double) * n);
  if (info) {
    info[0] = p->pd_minflux;
    info[1] = p->pd_dflux;
    info[2] = p->last_split;
    info[3] = p->last_peak;
    info[4] = p->sgpos;
    info[5] = p->sgneg;
    info[6] = p->last_integral;
    info[7] = p->pd_islog;
  }
  if (rc) return rc;
  if (pd_out) memcpy(pd_out, p->pd, sizeof(double) * n);
  return ORC_OK;
}

int orc_like_field_loglike(orc_like *L, int b, int f, double *out) {
  if (b < 0 || (unsigned)b >= L->nsets || !L->fac) return ORC_E_ARG;
  const orc_beamset *s = &L->sets[b];
  if (f < 0 || (unsigned)f >= s->nfields) return ORC_E_ARG;
  out[0] = orc_field_loglike_pd(L->fac, &s->fields[f], &out[1]);
  return ORC_OK;
}
