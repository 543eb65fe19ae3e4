/*
 * pofd_oracle.h -- CPU restatement of the reference 1D P(D) likelihood path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * HIP implementation in pofd_affine_amd/.  Only tests/, the smoke() entry
 * point and the cpu_baseline leg of bench.py may load it, and only as the
 * checker (or the reported CPU baseline), never as the product path.
 *
 * Every function restates one piece of the reference (aconley/pofd_affine,
 * GPL-3) and cites the file:line it follows.  The reference itself cannot be
 * compiled in this image (FFTW3, GSL, gtest and the autotools are absent), so
 * its two third-party numerical dependencies are restated here as well:
 *   - GSL (version unpinned by configure.ac:175-211): gsl_interp_cspline
 *     (natural cubic spline, interp/cspline.c: cspline_init/cspline_eval/
 *     coeff_calc) and gsl_linalg_solve_symm_tridiag (linalg/tridiag.c,
 *     LDL^T of a symmetric tridiagonal matrix), and gsl_interp_accel /
 *     gsl_interp_bsearch index semantics.
 *   - FFTW3 (version unpinned by configure.ac:102-140): r2c forward
 *     (e^{-2 pi i jk/N}) and unnormalised c2r backward transforms.  Restated
 *     as a radix-2 complex FFT of length N/2 plus the standard real
 *     pre/post-processing; FFTW's bit-level output is not reproducible, so the
 *     FFT is pinned against numpy's pocketfft in the tests instead.
 *
 * Build flags: -O2 -ffp-contract=off (the reference is built with g++ -O3 for
 * baseline x86-64, which has no FMA, so no contraction happens there either).
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   orc_cspline_*              scipy CubicSpline(bc_type='natural'), closed form
 *   orc_beam_*                 reference golden vectors test/beam_test.cc
 *   orc_knots_getR*            reference golden vectors test/model_test.cc,
 *                              power-law closed form, brute force
 *   orc_data_bin               mathematical bin-centre property + counts
 *   orc_rfft / orc_irfft       numpy.fft.rfft / irfft
 *   orc_pd_*                   reference golden vectors test/pofd_test.cc
 *   orc_like_* (end to end)    R == 0 Gaussian closed form, Campbell's theorem,
 *                              brute-force compound-Poisson P(D) at small n
 */
#ifndef POFD_ORACLE_H
#define POFD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; numerically identical to the POFD_* codes of the product
   header include/pofd_affine_amd.h (checked by a test). */
enum {
  ORC_OK = 0,
  ORC_E_SPLIT = 1,          /* PDFactory::findSplitPoint failure */
  ORC_E_INVALID_MODEL = 2,  /* numberCountsKnots::isValid() false in initPD */
  ORC_E_MEAN = 3,           /* non-finite mean in unwrapAndNormalizePD */
  ORC_E_BADLIKE = 4,        /* calcLikeSingle::bad_like returned */
  ORC_E_ARG = 5,            /* invalid argument (reference throws) */
  ORC_E_INTERNAL = 6
};

typedef struct orc_err {
  int code;
  char cls[48];
  char method[48];
  char msg[512];
} orc_err;

/* ---------------- natural cubic spline (GSL cspline) ---------------- */
/* c_out[0..n-1] receives the GSL "c" coefficients (half second derivative) */
int orc_cspline_init(size_t n, const double *x, const double *y, double *c_out);
double orc_cspline_eval(size_t n, const double *x, const double *y,
                        const double *c, double xv);
/* index i with x[i] <= xv < x[i+1] (gsl_interp_bsearch over [0, n-1]) */
size_t orc_bsearch(const double *x, double xv, size_t lo, size_t hi);

/* ---------------- beam (src/beam.cc) ---------------- */
typedef struct orc_beam {
  double pixsize, minval;
  uint32_t npos, nneg;
  int haspos, hasneg;
  double *pospix, *negpix, *posinv, *neginv;
  double totpos, totpossq, totneg, totnegsq;
  uint32_t nbins;
  int is_pos_hist, is_neg_hist;
  uint32_t posnbins, negnbins;
  double *posweights, *poshistval, *negweights, *neghistval;
} orc_beam;

orc_beam *orc_beam_new(void);
void orc_beam_free(orc_beam *bm);
/* beam::readFile after the FITS read: pix is the image in FITS row order */
int orc_beam_set(orc_beam *bm, const double *pix, size_t n, double pixsize,
                 double minval, orc_err *err);
int orc_beam_make_histogram(orc_beam *bm, uint32_t nbins, orc_err *err);
/* accessors for tests */
void orc_beam_info(const orc_beam *bm, double *out /* 16 */);
const double *orc_beam_array(const orc_beam *bm, int which);

/* ---------------- knot model (src/numberCountsKnots*.cc) ------------- */
#define ORC_MAXKNOTS 64
typedef struct orc_knots {
  unsigned nknots;
  int loaded;
  double knots[ORC_MAXKNOTS];
  double logknots[ORC_MAXKNOTS];
  double logknotvals[ORC_MAXKNOTS];
  double c[ORC_MAXKNOTS];
} orc_knots;

int orc_knots_set_positions(orc_knots *m, unsigned n, const double *pos,
                            orc_err *err);
int orc_knots_set_params(orc_knots *m, unsigned nparams, const float *p,
                         orc_err *err);
int orc_knots_is_valid(const orc_knots *m);
double orc_knots_number_counts(const orc_knots *m, double s);
void orc_knots_rrange(const orc_knots *m, const orc_beam *bm, double *minrf,
                      double *maxrf);
double orc_knots_getR1(const orc_knots *m, double f, const orc_beam *bm);
void orc_knots_getR(const orc_knots *m, unsigned n, const double *flux,
                    const orc_beam *bm, double *R);
double orc_knots_regularize(const orc_knots *m, double alpha);
/* splineInt (src/numberCountsKnotsSpline.cc:300-343): getNS (alpha 0),
   getFluxPerArea (1), getFluxSqPerArea (2) */
double orc_knots_splineint(const orc_knots *m, double alpha);
void orc_gl_nodes(unsigned n, double *x, double *w);

/* ---------------- data (src/fitsData.cc) ---------------- */
void orc_data_minmax(const double *d, size_t n, double *mn, double *mx);
double orc_data_mean_subtract(double *d, size_t n);
int orc_data_bin(const double *d, size_t n, uint32_t nbins, uint32_t *bins,
                 double *bincent0, double *bindelta);

/* ---------------- FFT (FFTW r2c / c2r semantics) ---------------- */
/* out: (n/2+1) interleaved complex */
int orc_rfft(unsigned n, const double *in, double *out);
/* in: (n/2+1) interleaved complex; out: n reals, unnormalised backward */
int orc_irfft(unsigned n, const double *in, double *out);
/* 2D, n x n real <-> n x (n/2+1) complex (FFTW multi-dim layout) */
int orc_rfft2(unsigned n, const double *in, double *out);
int orc_irfft2(unsigned n, const double *in, double *out);

/* ---------------- PD helpers (src/PD.cc) ---------------- */
double orc_pd_integral(unsigned n, const double *pd, double dflux, int islog);
double orc_pd_mean(unsigned n, const double *pd, double minflux, double dflux,
                   int islog, int donorm);
void orc_pd_mean_var(unsigned n, const double *pd, double minflux,
                     double dflux, int islog, int donorm, double *mean,
                     double *var);
double orc_pd_interp(unsigned n, const double *pd, double minflux,
                     double dflux, double x);
double orc_pd_loglike_binned(unsigned n, const double *pd, double minflux,
                             double dflux, int islog, const uint32_t *bins,
                             uint32_t nbins, double bincent0, double bindelta);
double orc_pd_loglike_unbinned(unsigned n, const double *pd, double minflux,
                               double dflux, int islog, const double *flux,
                               size_t ndata, double *compensated);

/* ---------------- likelihood context (calcLike + PDFactory) ---------- */
typedef struct orc_like orc_like;

orc_like *orc_like_new(unsigned fftsize, unsigned ninterp);
void orc_like_free(orc_like *L);
const orc_err *orc_like_error(const orc_like *L);
/* precise = 1: transforms and exponentials in x87 extended precision (the
   exact-math value of the reference's pipeline); 0: plain double as the
   reference computes it (default). */
void orc_like_set_precise(orc_like *L, int precise);
/* the reference's TIMING categories (src/PDFactory.cc:93-125) plus the
   likelihood sum and what none of them covers; thread CPU seconds summed
   over the batch workers */
enum { ORC_T_R = 0, ORC_T_P0, ORC_T_FFT, ORC_T_POS, ORC_T_SPLIT, ORC_T_COPY,
       ORC_T_NORM, ORC_T_MEAN, ORC_T_LOG, ORC_T_LIKE, ORC_T_OTHER,
       ORC_NTIME };
/* on != 0: the batch workers time their stages; resets the totals */
void orc_like_set_timing(orc_like *L, int on);
/* out[ORC_NTIME]: totals since the last orc_like_set_timing */
void orc_like_timing(const orc_like *L, double *out);
/* test hook: every split index found is moved by `shift` before the
   rotation (0 = the reference's behaviour) */
void orc_like_set_split_shift(orc_like *L, int shift);
/* PDFactory::findSplitPoint (src/PDFactory.cc:691-921) on a given array,
   with the factory members it reads passed in */
int orc_find_split_array(const double *pofd, unsigned n, double dflux,
                         double sgpos, double sgneg, double minflux_R,
                         unsigned wrapRidx, unsigned *out, double *peak,
                         orc_err *err);
int orc_like_set_knot_positions(orc_like *L, unsigned n, const double *pos);
/* adds a beam set (takes a copy of the beam); returns its index */
int orc_like_add_beamset(orc_like *L, const orc_beam *bm);
/* adds a dataset to beam set b.  bins==NULL -> unbinned (pix used). */
int orc_like_add_field(orc_like *L, int b, const double *pix, size_t npix,
                       const uint32_t *bins, uint32_t nbins, double bincent0,
                       double bindelta, double sigma_base, double like_norm);
/* calcLikeSingle::setSigmaBase / setExpConf / setRRange (per beam set) */
int orc_like_finalize_setup(orc_like *L, const float *theta, unsigned npar,
                            double exp_conf);
int orc_like_beamset_info(const orc_like *L, int b, double *out /* 8 */);
/* calcLike::getLogLike.  Returns status; *loglike set on success. */
int orc_like_loglike(orc_like *L, const float *theta, unsigned npar,
                     double *loglike);
/* batch over walkers with nthreads worker threads (CPU baseline) */
int orc_like_loglike_batch(orc_like *L, const float *theta, unsigned nwalk,
                           unsigned stride, unsigned npar, double *loglike,
                           int32_t *status, unsigned nthreads);

/* Stage-level access for parity tests: run initPD for beam set b at theta
   and return the intermediates.  Sizes: rinterp_flux/vals ninterp,
   rvals n, prtrans 2*(n/2+1).  scalars[8] = dflux, minflux_R, wrapRidx,
   mn, var_noi, doshift, shift, r0.  Any output pointer may be NULL. */
int orc_like_initpd(orc_like *L, int b, const float *theta, unsigned npar,
                    double *rinterp_flux, double *rinterp_vals, double *rvals,
                    double *prtrans, double *scalars);
/* RinterpFlux_neg / RinterpVals_neg of the last orc_like_initpd;
   ORC_E_ARG when the beam had no negative part */
int orc_like_rinterp_neg(orc_like *L, double *flux, double *vals);
/* getPD for field f of beam set b (after orc_like_initpd on the same
   context).  pd_out n values; info[8] = minflux, dflux, splitidx, peak,
   sgpos, sgneg, integral_before_norm, islog.  setlog as getPD(setLog). */
int orc_like_getpd(orc_like *L, int b, int f, double sigmult, int setlog,
                   double *pofd_raw, double *pd_out, double *info);
/* loglike of one field with the most recent getpd (setlog=1):
   out[0] = reference-order sum, out[1] = compensated sum */
int orc_like_field_loglike(orc_like *L, int b, int f, double *out);

#ifdef __cplusplus
}
#endif
#endif
