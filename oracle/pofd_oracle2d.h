/*
 * pofd_oracle2d.h -- CPU restatement of the reference 2D (two-band) P(D)
 * likelihood path.
 *
 * TEST INFRASTRUCTURE ONLY, like pofd_oracle.h: the parity checker for the
 * HIP 2D path, loaded only by tests/, smoke() and bench.py's cpu_baseline.
 *
 * Restates (aconley/pofd_affine, GPL-3):
 *   doublebeam                   src/doublebeam.cc:331-756
 *   numberCountsDoubleLogNormal  src/numberCountsDoubleLogNormal.cc:
 *                                1010-1052 (setParams), 1060-1130 (validity),
 *                                1172-1200 (sigma/offset), 1230-1255 (counts),
 *                                1352-1480 (max flux, log-normal solver),
 *                                1481-1547 (R range), 1682-2007 (getR)
 *   fitsDataDouble               src/fitsDataDouble.cc:481-608
 *   PDDouble                     src/PDDouble.cc:142-230, 355-452, 976-1222
 *   PDFactoryDouble              src/PDFactoryDouble.cc:18-48, 313-722,
 *                                784-1118, 1145-1460
 *   calcLikeDoubleSingle/calcLikeDouble
 *                                src/calcLikeDouble.cc:360-420, 496-620,
 *                                1093-1097, 1195-1275
 * plus GSL's gsl_interp_linear evaluation and the Brent root solver
 * (roots/brent.c) with gsl_root_test_interval, restated from the published
 * algorithms (GSL version unpinned by configure.ac:175-211).
 */
#ifndef POFD_ORACLE2D_H
#define POFD_ORACLE2D_H

#include <stddef.h>
#include <stdint.h>

#include "pofd_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* PDFactoryDouble::findSplitPoint (src/PDFactoryDouble.cc:784-975) on a
   given marginal of length cs */
int orc2_find_split_array(const double *data, unsigned cs, double dflux,
                          double sigmapos, double sigmaneg, unsigned *out,
                          orc_err *err);

/* ---------------- doublebeam ---------------- */
typedef struct orc_dbeam {
  double pixsize, minval;
  int hassign[4];
  uint32_t npix[4];
  double *pix1[4], *pix2[4], *ipix1[4], *ipix2[4], *logratio[4];
  double tot1[4], tot2[4], totsq1[4], totsq2[4];
  double min1[4], max1[4], min2[4], max2[4];
  uint32_t nbins;
  int ishist[4];
  uint32_t nhist[4];
  double *binw[4], *binv1[4], *binv2[4], *binlogratio[4];
} orc_dbeam;

orc_dbeam *orc_dbeam_new(void);
void orc_dbeam_free(orc_dbeam *bm);
/* doublebeam::setBeams: b1, b2 the two images in FITS row order */
int orc_dbeam_set(orc_dbeam *bm, const double *b1, const double *b2, size_t n,
                  double pixsize, double minval, orc_err *err);
int orc_dbeam_make_histogram(orc_dbeam *bm, uint32_t nbins, orc_err *err);
/* out[64]: per sign s (0 pp, 1 pn, 2 np, 3 nn) at 12*s: hassign, npix,
   ishist, nhist, tot1, tot2, totsq1, totsq2, min1, max1, min2, max2;
   out[48..50] = pixsize, minval, nbins */
void orc_dbeam_info(const orc_dbeam *bm, double *out);
/* which: 0 pix1, 1 pix2, 2 ipix1, 3 ipix2, 4 logratio, 5 binw, 6 binv1,
   7 binv2, 8 binlogratio */
const double *orc_dbeam_array(const orc_dbeam *bm, int sgn, int which,
                              uint32_t *n);

/* ---------------- numberCountsDoubleLogNormal ---------------- */
#define ORC_MAXK2 64
typedef struct orc_lnmodel {
  unsigned nknots, nsigma, noffset;
  double knots[ORC_MAXK2], logknots[ORC_MAXK2], logknotvals[ORC_MAXK2];
  double kc[ORC_MAXK2]; /* band-1 cspline c coefficients (log2 space) */
  double sigmaknots[ORC_MAXK2], sigmavals[ORC_MAXK2], sc[ORC_MAXK2];
  double offsetknots[ORC_MAXK2], offsetvals[ORC_MAXK2], oc[ORC_MAXK2];
  double min_sigma;
  int pos_loaded, vals_loaded;
  int knots_valid, sigmas_valid, offsets_valid;
} orc_lnmodel;

void orc_ln_init(orc_lnmodel *m);
int orc_ln_set_positions(orc_lnmodel *m, unsigned nk, const double *kp,
                         unsigned ns, const double *sp, unsigned no,
                         const double *op, orc_err *err);
int orc_ln_set_min_sigma(orc_lnmodel *m, double mins, orc_err *err);
int orc_ln_set_params(orc_lnmodel *m, unsigned npar, const float *F,
                      orc_err *err);
int orc_ln_is_valid(const orc_lnmodel *m);
double orc_ln_sigma(const orc_lnmodel *m, double f1);
double orc_ln_offset(const orc_lnmodel *m, double f1);
double orc_ln_counts(const orc_lnmodel *m, double f1, double f2);
double orc_ln_band1_counts(const orc_lnmodel *m, double f1);
/* splineInt(alpha, beta) (src/numberCountsDoubleLogNormal.cc:1549-1673):
   getNS (0,0), getFluxPerArea band 1 (1,0) / band 2 (0,1), getFluxSqPerArea
   band 1 (2,0) / band 2 (0,2) */
double orc_ln_splineint(const orc_lnmodel *m, double alpha, double beta);
/* getMaxFlux: out[2] = band 1, band 2 */
int orc_ln_maxflux(const orc_lnmodel *m, double *out, orc_err *err);
/* getRRange: out[4] = min1, max1, min2, max2 */
int orc_ln_rrange(const orc_lnmodel *m, const orc_dbeam *bm, double *out,
                  orc_err *err);
double orc_ln_getR1(const orc_lnmodel *m, double f1, double f2,
                    const orc_dbeam *bm);
/* array getR: R[n1*n2] row-major, including the pixel-area factor */
void orc_ln_getR(const orc_lnmodel *m, unsigned n1, const double *f1,
                 unsigned n2, const double *f2, const orc_dbeam *bm,
                 double *R);

/* ---------------- fitsDataDouble ---------------- */
void orc_data2_minmax(const double *d1, const double *d2, size_t n,
                      double *out /* 4 */);
void orc_data2_mean_subtract(double *d1, double *d2, size_t n,
                             double *mean /* 2 */);
/* bins[nb1*nb2]; out[4] = bincent01, bindelta1, bincent02, bindelta2 */
int orc_data2_bin(const double *d1, const double *d2, size_t n, uint32_t nb1,
                  uint32_t nb2, uint32_t *bins, double *out);

/* ---------------- PDDouble ---------------- */
double orc_pd2_integral(unsigned n1, unsigned n2, const double *pd,
                        double dflux1, double dflux2, int islog);
void orc_pd2_means(unsigned n1, unsigned n2, const double *pd, double minflux1,
                   double dflux1, double minflux2, double dflux2, int islog,
                   int donorm, double *out /* 2 */);
double orc_pd2_loglike_unbinned(unsigned n1, unsigned n2, const double *pd,
                                double minflux1, double dflux1,
                                double minflux2, double dflux2, int islog,
                                const double *d1, const double *d2,
                                size_t ndata);
double orc_pd2_loglike_binned(unsigned n1, unsigned n2, const double *pd,
                              double minflux1, double dflux1, double minflux2,
                              double dflux2, int islog, const uint32_t *bins,
                              uint32_t nb1, uint32_t nb2, double bincent01,
                              double bindelta1, double bincent02,
                              double bindelta2);

/* ---------------- likelihood context (calcLikeDouble) ---------------- */
typedef struct orc_like2 orc_like2;

orc_like2 *orc_like2_new(unsigned fftsize, unsigned nedge, int edge_integ);
void orc_like2_free(orc_like2 *L);
const orc_err *orc_like2_error(const orc_like2 *L);
int orc_like2_set_positions(orc_like2 *L, unsigned nk, const double *kp,
                            unsigned ns, const double *sp, unsigned no,
                            const double *op, double min_sigma);
int orc_like2_add_beamset(orc_like2 *L, const orc_dbeam *bm);
/* bins == NULL: unbinned pixel pairs (d1, d2) */
int orc_like2_add_field(orc_like2 *L, int b, const double *d1,
                        const double *d2, size_t npix, const uint32_t *bins,
                        uint32_t nb1, uint32_t nb2, const double *bininfo,
                        double sigma_base1, double sigma_base2,
                        double like_norm);
/* setSigmaBase1/2 (like_offset), exp_conf, setRRanges(theta) */
int orc_like2_finalize_setup(orc_like2 *L, const float *theta, unsigned npar,
                             double exp_conf1, double exp_conf2);
int orc_like2_set_rrange(orc_like2 *L, int b, const double *range /* 4 */);
/* out[12]: minR1, maxR1, minR2, maxR2, maxsig1, maxsig2, mind1, maxd1,
   mind2, maxd2, nfields, like_offset of field 0 */
int orc_like2_beamset_info(const orc_like2 *L, int b, double *out);
int orc_like2_loglike(orc_like2 *L, const float *theta, unsigned npar,
                      double *loglike);
int orc_like2_loglike_batch(orc_like2 *L, const float *theta, unsigned nwalk,
                            unsigned stride, unsigned npar, double *loglike,
                            int32_t *status, unsigned nthreads);
/* initPD for beam set b at theta; scalars[12] = dflux1, dflux2, mn1, mn2,
   var1, var2, shift1, shift2, doshift1, doshift2, r0, wrap1.  rvals
   (n*n, R times dflux1*dflux2) and prtrans (2*n*(n/2+1)) may be NULL. */
int orc_like2_initpd(orc_like2 *L, int b, const float *theta, unsigned npar,
                     double *rvals, double *prtrans, double *scalars);
/* getPD for field f (after initpd): pd_out n*n (rotated, normalised, log2
   when setlog), info[8] = minflux1, minflux2, dflux1, dflux2, split1,
   split2, sgpos1, sgpos2 */
int orc_like2_getpd(orc_like2 *L, int b, int f, double sigmult1,
                    double sigmult2, int setlog, double *pd_out, double *info);

#ifdef __cplusplus
}
#endif
#endif
