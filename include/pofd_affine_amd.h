/*
 * pofd_affine_amd.h -- C ABI of the MI355X (gfx950) P(D) likelihood library.
 *
 * This is the drop-in boundary for the likelihood call that the reference's
 * MPI slaves make (affineEnsemble::slaveSample -> pofdMCMC::getLogLike ->
 * calcLike::getLogLike, src/affineEnsemble.cc:1490, src/pofdMCMC.cc:347-352,
 * include/calcLike.h:275).  The reference has no foreign-function layer of its
 * own; each entry point below names the reference interface it replaces.
 * Plain C types only: pointers, sizes, doubles, float32 parameter vectors
 * (paramSet stores float, include/paramSet.h:20).
 *
 * Threading: a context is bound to one host thread and one HIP device, like
 * the reference's one-calcLike-per-MPI-rank model (calcLike.h:47-48 mutable
 * state is not re-entrant).  All functions return POFD_OK (0) or a POFD_E_*
 * code; pofd_last_error() then holds "class::method: message" in the style of
 * affineExcept (include/affineExcept.h).
 */
#ifndef POFD_AFFINE_AMD_H
#define POFD_AFFINE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes, also used per walker in the batch outputs. */
enum {
  POFD_OK = 0,
  POFD_E_SPLIT = 1,         /* PDFactory::findSplitPoint threw
                               (src/PDFactory.cc:753-920) */
  POFD_E_INVALID_MODEL = 2, /* initPD "Invalid model": a knot value is NaN
                               (src/numberCountsKnots.cc:273-284,
                               src/PDFactory.cc:1046-1047) */
  POFD_E_MEAN = 3,          /* unwrapAndNormalizePD non-finite mean
                               (src/PDFactory.cc:998-1004) */
  POFD_E_BADLIKE = 4,       /* calcLikeSingle::bad_like returned: sigma
                               multiplier < 0 (src/calcLike.cc:403); the
                               walker's loglike carries 1e25 per beam set as
                               in the reference */
  POFD_E_ARG = 5,           /* invalid argument (reference throws) */
  POFD_E_INTERNAL = 6,      /* HIP runtime failure or unsupported setup */
  POFD_E_NOGPU = 7          /* no usable gfx950 device */
};

/* Options for pofd_set_option. */
enum {
  /* Spectral components whose Gaussian noise factor
     exp(-sigma^2 omega^2 / 2) is below exp(-value) are treated as zero.
     Default 44.3614195558365 (= 64 ln 2, i.e. below 2^-64 of the DC term).
     746 reproduces exactly the components the reference's exp() already
     underflows to 0.0 (src/PDFactory.cc:1180-1183). */
  POFD_OPT_SPECTRAL_CUT = 1,
  /* Nonzero: the batched entry points record HIP events around each
     kernel stage on the launch stream (read back with
     pofd1d_stage_times); 0, the default, records none.  No reference
     counterpart: the reference's TIMING build keeps per-stage wall-clock
     totals in PDFactory (src/PDFactory.cc:93-126). */
  POFD_OPT_STAGE_TIMING = 2
};

/* Kernel stages of one beam set's batch, in launch order. */
enum {
  POFD_STAGE_PREP = 0,      /* knot spline, R at the nodes, node spline */
  POFD_STAGE_FORWARD = 1,   /* R on the grid, forward transform, spectrum */
  POFD_STAGE_INVERSE = 2,   /* per dataset: inverse transform, P(D) */
  POFD_STAGE_FINALIZE = 3,  /* per dataset: split, normalise, binned logL */
  POFD_STAGE_PIXELS = 4,    /* unbinned logL over the pixels */
  POFD_STAGE_COMBINE = 5,   /* sum over datasets into the outputs */
  POFD_NSTAGES = 6
};

typedef struct pofd_ctx pofd_ctx;
typedef struct pofd_beam pofd_beam;

/* ---- context --------------------------------------------------------- */
int pofd_create(pofd_ctx **out, int hip_device);
void pofd_destroy(pofd_ctx *ctx);
const char *pofd_last_error(const pofd_ctx *ctx);
int pofd_set_option(pofd_ctx *ctx, int option, double value);

/* ---- beams: replaces beam::readFile (after the FITS read) and
        beam::makeHistogram, src/beam.cc:295-458, 467-556.  Host code,
        bit-exact with the reference (glibc log2, same operation order). ---- */
int pofd_beam_create(pofd_beam **out, const double *pix, uint64_t npix,
                     double pixsize_arcsec, double minval);
int pofd_beam_make_histogram(pofd_beam *bm, uint32_t nbins);
void pofd_beam_destroy(pofd_beam *bm);
/* which: 0 pos pixels (sorted), 1 |neg| pixels (sorted), 2 pos inverse,
   3 neg inverse, 4 pos hist weights, 5 pos hist inverse values,
   6 neg hist weights, 7 neg hist inverse values.  *n receives the length;
   out may be NULL to query it. */
int pofd_beam_get(const pofd_beam *bm, int which, double *out, uint32_t *n);
/* info[16]: npos, nneg, haspos, hasneg, pixsize, minval, is_pos_hist,
   is_neg_hist, posnbins, negnbins, area (deg^2), area_pos, area_neg,
   area_pix, area_sq, max positive pixel */
int pofd_beam_info(const pofd_beam *bm, double *info);

/* ---- FITS ingestion: the read side of beam::readFile and
        fitsData::readData (src/beam.cc:295-380, src/fitsData.cc:56-315).
        IMAGE HDUs with BITPIX 8/16/32/64/-32/-64 and BZERO/BSCALE; "file[k]"
        selects HDU k; maps use the image/IMAGE extension (else the primary)
        and drop pixels whose mask/MASK value is non-zero.  Call with a NULL
        buffer to query the size.  pofd_fits_last_error() describes the last
        failure of the calling thread. */
int pofd_fits_read_beam(const char *path, double *pix, uint64_t capacity,
                        uint64_t *npix, double *pixsize_arcsec);
int pofd_fits_read_map(const char *path, int ignore_mask, double *data,
                       uint64_t capacity, uint64_t *n);
/* Pixel pairs of a band-1 and a band-2 map (fitsDataDouble::readData,
   src/fitsDataDouble.cc:305-406): with both masks a pixel is kept when
   (mask1 & mask2) == 0, with one mask when it is 0.  NULL buffers query n. */
int pofd_fits_read_map_pair(const char *path1, const char *path2,
                            int ignore_mask, double *d1, double *d2,
                            uint64_t capacity, uint64_t *n);
const char *pofd_fits_last_error(void);

/* ---- data: replaces fitsData::applyBinning and fitsData::meanSubtract,
        src/fitsData.cc:393-437.  Host code, bit-exact. ---- */
int pofd_data_bin(const double *data, uint64_t n, uint32_t nbins,
                  uint32_t *bins_out, double *bincent0, double *bindelta);
int pofd_data_mean_subtract(double *data, uint64_t n, double *mean_out);

/* ---- 1D likelihood setup: replaces calcLike::readDataFromFiles grouping
        (src/calcLike.cc:688-758), calcLikeSingle::setSigmaBase /
        setLikeNorm / setExpConf (src/calcLike.cc:241-308) and
        calcLike::setKnotPositions / setFFTSize / setNInterp. ---- */
/* knots in Jy (double, strictly increasing, > 0); nflux a power of two
   from 2^13 to 2^18 (the reference plans FFTW for any size; the pruned
   four-step transform here needs 256 x M2 complex points, M2 a power of
   two from 16 to 512); ninterp in [16, 2048] */
int pofd1d_configure(pofd_ctx *ctx, uint32_t nflux, uint32_t ninterp,
                     const double *knot_pos, uint32_t nknots,
                     double exp_conf);
/* adds a beam set (the beam is copied); *beamset_id receives its index */
int pofd1d_add_beamset(pofd_ctx *ctx, const pofd_beam *bm, int *beamset_id);
/* adds a dataset to a beam set.  Unbinned: pix/npix (bins NULL).
   Binned: bins/nbins/bincent0/bindelta from pofd_data_bin and npix = the
   number of pixels that were binned (used for like_offset). */
int pofd1d_add_field(pofd_ctx *ctx, int beamset_id, const double *pix,
                     uint64_t npix, const uint32_t *bins, uint32_t nbins,
                     double bincent0, double bindelta, double sigma_base,
                     double like_norm, int *field_id);
/* Builds the theta-independent device tables.  Replaces calcLike::setRRanges
   (src/calcLike.cc:827-831 -> calcLikeSingle::setRRange :346-371); the R
   range depends only on knot positions, beam and data. */
int pofd1d_prepare(pofd_ctx *ctx);
/* Overrides the R range of a beam set (call after pofd1d_prepare). */
int pofd1d_set_rrange(pofd_ctx *ctx, int beamset_id, double min_rflux,
                      double max_rflux);
/* out[8]: minRFlux, maxRFlux, maxsigma_base, minDataFlux, maxDataFlux,
   nfields, exp_conf, dflux */
int pofd1d_beamset_info(const pofd_ctx *ctx, int beamset_id, double *out);

/* ---- hot calls: replace calcLike::getLogLike(const paramSet&, bool&)
        (src/calcLike.cc:884-955) for a batch of walkers.  theta rows use the
        pofdMCMC layout [knot log10 values..., sigma_mult, bonus...]
        (src/pofdMCMC.cc:67-91); row stride theta_stride floats. ---- */
/* host buffers; synchronous */
int pofd1d_loglike_batch(pofd_ctx *ctx, const float *theta, uint32_t n_walkers,
                         uint32_t theta_stride, double *loglike_out,
                         int32_t *status_out);
/* device buffers on the context's device; asynchronous on `stream`
   (a hipStream_t, 0 = the null stream) */
int pofd1d_loglike_batch_device(pofd_ctx *ctx, const float *d_theta,
                                uint32_t n_walkers, uint32_t theta_stride,
                                double *d_loglike_out, int32_t *d_status_out,
                                void *stream);
/* As pofd1d_loglike_batch_device for the first *d_count rows (an int32 in
   device memory, 0 <= *d_count <= max_walkers, read when the kernels run):
   a caller whose batch size is produced on the GPU (the delayed-acceptance
   sampler's first-stage survivors) launches without reading it back.  No
   reference counterpart.  Outputs past *d_count are left unwritten. */
int pofd1d_loglike_batch_device_count(pofd_ctx *ctx, const float *d_theta,
                                      uint32_t max_walkers,
                                      const int32_t *d_count,
                                      uint32_t theta_stride,
                                      double *d_loglike_out,
                                      int32_t *d_status_out, void *stream);
/* No counterpart in the reference.  For callers that cannot allocate device
   memory -- code under HIP graph capture, tight serving loops, harnesses
   that forbid hipMalloc: pofd1d_loglike_batch_device with the per-walker
   workspace of every beam set laid out in a fixed device scratch of 512 MiB
   per GPU (reserved when the library's code is loaded, held until the
   process ends) instead of memory the context allocates; it calls no
   hipMalloc or hipFree and records no stage events.  A batch whose
   workspace exceeds the scratch returns POFD_E_ARG and says how much it
   needs (at nflux 2^17 with one dataset: up to about 64 walkers).  All
   calls on one GPU share its scratch: calls that overlap on one GPU (other
   streams, other host threads, other contexts) race.  Results equal
   pofd1d_loglike_batch_device's bit for bit; the debug accessors describe
   the allocating entry points' batches only. */
int pofd1d_loglike_batch_device_noalloc(pofd_ctx *ctx, const float *d_theta,
                                        uint32_t n_walkers,
                                        uint32_t theta_stride,
                                        double *d_loglike_out,
                                        int32_t *d_status_out, void *stream);
/* Priors and regularisation of calcLike::getLogLike (src/calcLike.cc:
   923-952), added on the device to every batch's log likelihood after the
   dataset sum (walkers whose datasets failed are left as they are):
   POFD_PRIOR_SIGMA   -1/2 ln 2pi - ln w - ((sigma_mult - 1)/w)^2 / 2
   POFD_PRIOR_CFIRB   the same in (cfirb_mean - int S n dS) / cfirb_sigma
   POFD_PRIOR_POISSON the same in (poisson_mean - int S^2 n dS) / poisson_sigma
   and, when reg_alpha > 0, differenceRegularize(reg_alpha)
   (src/numberCountsKnotsSpline.cc:540-559).  flags 0 and reg_alpha 0 turn
   them off (the default). */
enum { POFD_PRIOR_SIGMA = 1, POFD_PRIOR_CFIRB = 2, POFD_PRIOR_POISSON = 4 };
int pofd1d_set_priors(pofd_ctx *ctx, int flags, double sigma_width,
                      double cfirb_mean, double cfirb_sigma,
                      double poisson_mean, double poisson_sigma,
                      double reg_alpha);
/* Several fields with their own nflux (one context each) sharing theta, as
   one captured HIP graph, no reference counterpart: the contexts' batches
   run in order on `stream`, each adding its log likelihood (and keeping an
   earlier failure's status) to d_loglike_out, so one pofd1d_graph_launch
   evaluates the sum over all the fields.  The workspaces are sized before
   the capture; d_theta, d_loglike_out and d_status_out are bound to the
   graph (write new theta rows into d_theta between launches).  Stage
   timing must be off.  Priors set on any of the contexts are added once by
   that context.
   The graph holds the contexts' device pointers and launch parameters as
   they were at the capture.  Anything that replaces them afterwards -- an
   eager batch larger than any before (the workspace grows), prepare or
   set_rrange (tables uploaded again), configure, set_option,
   set_parts, set_priors, autotune_parts -- makes pofd1d_graph_launch
   return POFD_E_ARG (message on that context) without launching; create
   the graph again.  The contexts must outlive the graph. */
typedef struct pofd_graph pofd_graph;
int pofd1d_graph_create(pofd_ctx **ctxs, uint32_t n_ctx, const float *d_theta,
                        uint32_t n_walkers, uint32_t theta_stride,
                        double *d_loglike_out, int32_t *d_status_out,
                        void *stream, pofd_graph **out);
int pofd1d_graph_launch(pofd_graph *graph, void *stream);
void pofd1d_graph_destroy(pofd_graph *graph);
/* Tuning, no reference counterpart: workgroups per walker of the forward
   and inverse transforms.  0 (the default) chooses by batch size,
   min(4, 256 / n_walkers) and at least 1, with 3 taken as 2, so that small
   batches (such as the survivors of a delayed-acceptance first stage)
   still fill the GPU's 256 compute units; 1, 2 or 4 forces it (halved
   while walkers x parts exceeds 256); other values return POFD_E_ARG.
   Every sum the split touches runs over four fixed column groups combined
   as (G0 + G1) + (G2 + G3), so the results are the same bits for every
   part count and batch size. */
int pofd1d_set_parts(pofd_ctx *ctx, int parts);
/* Autotuning, no reference counterpart: measures the workgroups per walker.
   For each of the n_sizes batch sizes (1 to 256) every allowed split (1, 2
   or 4, at most 256 / size) is timed with HIP events over `reps` batches of the
   first `size` rows of theta (host memory, at least max(sizes) rows of
   `theta_stride` floats) after one warm-up batch, and the fastest is
   written to chosen[i].  The context then uses the measured choice for
   every batch size (an unmeasured size takes the choice of the nearest
   measured one) whenever pofd1d_set_parts is 0, in host-count and
   device-count batches alike.  n_sizes = 0 returns to the size rule.
   Stage timing must be off. */
int pofd1d_autotune_parts(pofd_ctx *ctx, const float *theta, uint32_t theta_stride,
                          const int32_t *sizes, uint32_t n_sizes, uint32_t reps,
                          int32_t *chosen);

/* ---- log likelihood of data given a caller's P(D) ---- */
/* Replaces double PD::getLogLike(const fitsData&) const (include/PD.h:86,
   src/PD.cc:538-677): the P(D) pd[n] on the grid minflux + i*dflux (log2
   values when is_log, as PD::logflat), linearly interpolated at every
   pixel (pix/npix; edges x <= minflux, x >= maxflux - dflux) or at every
   bin centre bincent0 + i*bindelta weighted by bins[i] (edges < and >),
   summed and converted to natural log.  Host buffers in, *loglike_out out;
   runs on ctx's device (no other context state is used).  Errors as the
   reference throws them: no P(D), no data. */
int pofd_pd_loglike(pofd_ctx *ctx, const double *pd, uint32_t n, double minflux,
                    double dflux, int is_log, const double *pix, uint64_t npix,
                    const uint32_t *bins, uint32_t nbins, double bincent0,
                    double bindelta, double *loglike_out);
/* Replaces double PDDouble::getLogLike(const fitsDataDouble&) const
   (include/PDDouble.h:95, src/PDDouble.cc:976-1222): pd is n1 x n2 with the
   band-1 index outer (PDDouble::pd_); bilinear interpolation at the pixel
   pairs (pix1, pix2) or at the nbins1 x nbins2 histogram (band-1 bin
   index outer) with the reference's edge tests. */
int pofd_pd2_loglike(pofd_ctx *ctx, const double *pd, uint32_t n1, uint32_t n2,
                     double minflux1, double dflux1, double minflux2,
                     double dflux2, int is_log, const double *pix1,
                     const double *pix2, uint64_t npix, const uint32_t *bins,
                     uint32_t nbins1, uint32_t nbins2, double bincent01,
                     double bindelta1, double bincent02, double bindelta2,
                     double *loglike_out);

/* ---- parity / debug accessors for the most recent batch ---- */
/* The P(D) as PDFactory::getPD(sigma, pd, setLog) leaves PD::pd_ (rotated,
   normalised, log2 when is_log > 0) plus pd.minflux / pd.dflux
   (src/PDFactory.cc:934-1009, src/PD.cc:169-184).  is_log < 0 returns the
   clamped c2r output before the split (PDFactory's internal pofd array). */
int pofd1d_get_pd(pofd_ctx *ctx, uint32_t walker, int beamset_id,
                  int field_id, int is_log, double *pd_out, double *minflux,
                  double *dflux);
/* info[32]: split index, peak, minidx, minval, jitter, integral (before
   normalisation), mean shift (-minflux), sgpos, sgneg, mn, var_noi, r0,
   kmax (last spectral index kept), nrows, field loglike (before offset),
   status, 1/integral, minflux, P(D) stored as log2, kmax of the beam set,
   sum P, sum j P, sum of P above the split, sigma, then wall-clock ticks
   (100 MHz) inside the kernels: prep, (unused), forward transform,
   spectrum and noisy rows, inverse transform, split search,
   normalise+logL, sum of these */
int pofd1d_get_debug(pofd_ctx *ctx, uint32_t walker, int beamset_id,
                     int field_id, double *info);
/* prtrans after initPD for k = 0..kmax (interleaved re, im), kmax from
   pofd1d_get_debug; *n receives kmax+1 */
int pofd1d_get_spectrum(pofd_ctx *ctx, uint32_t walker, int beamset_id,
                        double *out, uint32_t capacity, uint32_t *n);
/* R interpolation nodes and log2 R values (PDFactory::initRInterp,
   src/PDFactory.cc:426-527) of the most recent batch */
int pofd1d_get_rinterp(pofd_ctx *ctx, uint32_t walker, int beamset_id,
                       double *flux_out, double *log2r_out);
/* The same for the negative beam: RinterpFlux_neg / RinterpVals_neg
   (src/PDFactory.cc:486-524), ascending fluxes below zero; POFD_E_ARG when
   the beam has no negative pixels */
int pofd1d_get_rinterp_neg(pofd_ctx *ctx, uint32_t walker, int beamset_id,
                           double *flux_out, double *log2r_out);
/* Bonus quantities of the last batch, per walker w: out[3w..3w+2] =
   getNS, getFluxPerArea, getFluxSqPerArea of the knot model at that
   walker's theta (numberCountsKnotsSpline::splineInt,
   src/numberCountsKnotsSpline.cc:300-343; calcLike::getLogLike computes the
   last two for the CFIRB and Poisson priors and fillBonusParams,
   src/calcLike.cc:932-948, 958-980).  Integrated to double precision (the
   reference's GSL QAG-41 stops at relative 1e-5); NaN for an invalid
   model. */
int pofd1d_get_bonus(pofd_ctx *ctx, uint32_t nwalk, double *out);
/* With POFD_OPT_STAGE_TIMING on: waits for the recorded batches, writes the
   summed GPU milliseconds of each stage (ms_out[POFD_NSTAGES]) over every
   beam-set launch since the previous call, *nlaunch the number of such
   launches, and starts a new tally.  Up to 4096 launches are recorded per
   tally.  Two-band launches (pofd2d_loglike_batch*) enter the same tally
   with the stages: prep = model set-up; forward = R on the grid and its
   edges, moments, forward transforms and spectrum; inverse = noise factor
   and inverse transforms; finalize = split, rotation, normalisation;
   pixels = log likelihood; combine.  No reference counterpart (see
   POFD_OPT_STAGE_TIMING). */
int pofd1d_stage_times(pofd_ctx *ctx, double *ms_out, uint32_t *nlaunch);

/* ======================================================================
   Two-band (2D) path: calcLikeDouble / PDFactoryDouble / PDDouble with the
   numberCountsDoubleLogNormal colour model.
   ====================================================================== */
typedef struct pofd_dbeam pofd_dbeam;

/* ---- doublebeam::setBeams and makeHistogram, src/doublebeam.cc:331-584.
        Host code, bit-exact.  beam1/beam2 are the two images in FITS
        order, n pixels each; pixels with |b| > minval in both are kept and
        sorted into the pp, pn, np, nn sign components. ---- */
int pofd_dbeam_create(pofd_dbeam **out, const double *beam1,
                      const double *beam2, uint64_t n, double pixsize_arcsec,
                      double minval);
int pofd_dbeam_make_histogram(pofd_dbeam *bm, uint32_t nbins);
void pofd_dbeam_destroy(pofd_dbeam *bm);
const char *pofd_dbeam_last_error(void);
/* out[64]: per sign s at 12*s: hassign, npix, ishist, nhist, tot1, tot2,
   totsq1, totsq2, min1, max1, min2, max2; out[48..50] = pixsize, minval,
   nbins (doublebeam accessors, src/doublebeam.cc:598-720) */
int pofd_dbeam_info(const pofd_dbeam *bm, double *out);
/* which: 0 pix1, 1 pix2, 2 1/pix1, 3 1/pix2, 4 log(pix1/pix2), 5 bin
   weights, 6 bin 1/beam1, 7 bin 1/beam2, 8 bin log ratio; *n gets the
   length, out may be NULL (src/doublebeam.cc:722-776) */
int pofd_dbeam_get(const pofd_dbeam *bm, int sgn, int which, double *out,
                   uint32_t *n);

/* ---- fitsDataDouble::applyBinning and meanSubtract,
        src/fitsDataDouble.cc:517-608.  bins[nb1*nb2] row-major in band 1;
        info[4] = bincent01, bindelta1, bincent02, bindelta2. ---- */
int pofd_data2_bin(const double *d1, const double *d2, uint64_t n,
                   uint32_t nb1, uint32_t nb2, uint32_t *bins, double *info);
int pofd_data2_mean_subtract(double *d1, double *d2, uint64_t n,
                             double *mean);

/* ---- context set-up: calcLikeDouble(FFTSIZE, NEDGE, EDGEINTEG)
        (include/calcLikeDouble.h:241-243), setPositions of the model
        (src/numberCountsDoubleLogNormal.cc:632-920), setMinSigma (:792),
        setExpConf1/2.  n a power of two in [64, 4096]. ---- */
int pofd2d_configure(pofd_ctx *ctx, uint32_t n, uint32_t nedge,
                     int edge_integ, const double *knots, uint32_t nknots,
                     const double *sigmapos, uint32_t nsigma,
                     const double *offsetpos, uint32_t noffset,
                     double min_sigma, double exp_conf1, double exp_conf2);
/* calcLikeDouble::readDataFromFiles grouping by beam
   (src/calcLikeDouble.cc:930-1020): one beam set per beam pair */
int pofd2d_add_beamset(pofd_ctx *ctx, const pofd_dbeam *bm, int *beamset_id);
/* one dataset (pixel pairs, or a 2D histogram when bins != NULL) with its
   instrument noise per band and likelihood normalisation */
int pofd2d_add_field(pofd_ctx *ctx, int beamset_id, const double *d1,
                     const double *d2, uint64_t npix, const uint32_t *bins,
                     uint32_t nb1, uint32_t nb2, const double *bininfo,
                     double sigma_base1, double sigma_base2,
                     double like_norm, int *field_id);
/* calcLikeDouble::setRRanges(p) (src/calcLikeDouble.cc:1093-1097): R
   ranges of every beam set from the model at theta, like offsets, and the
   device tables */
int pofd2d_prepare(pofd_ctx *ctx, const float *theta, uint32_t npar);
/* override of a beam set's R range: range[4] = min1, max1, min2, max2 */
int pofd2d_set_rrange(pofd_ctx *ctx, int beamset_id, const double *range);
/* out[14]: minR1, maxR1, minR2, maxR2, maxsig1, maxsig2, min/max data band
   1, min/max data band 2, nfields, like_offset of dataset 0, dflux1,
   dflux2 */
int pofd2d_beamset_info(pofd_ctx *ctx, int beamset_id, double *out);

/* ---- hot call: calcLikeDouble::getLogLike (src/calcLikeDouble.cc:
        1195-1226, priors excluded) for nwalk walkers.  theta rows are the
        reference's float paramSet: [knot values (log10), sigma values,
        offset values, sigmult1, sigmult2, ...].  Status as the 1D path;
        a model failing isValid() gives POFD_E_INVALID_MODEL. ---- */
int pofd2d_loglike_batch(pofd_ctx *ctx, const float *theta, uint32_t nwalk,
                         uint32_t theta_stride, double *loglike_out,
                         int32_t *status_out);
int pofd2d_loglike_batch_device(pofd_ctx *ctx, const float *d_theta,
                                uint32_t nwalk, uint32_t theta_stride,
                                double *d_loglike, int32_t *d_status,
                                void *hip_stream);

/* Bonus quantities of the last batch, per walker w: out[5w..5w+4] = getNS,
   getFluxPerArea(0), getFluxPerArea(1), getFluxSqPerArea(0),
   getFluxSqPerArea(1) of the colour model at that walker's theta
   (numberCountsDoubleLogNormal::splineInt, src/numberCountsDoubleLogNormal.cc:
   1549-1673; used by calcLikeDouble's CFIRB/Poisson priors and
   fillBonusParams, src/calcLikeDouble.cc:1244-1316).  Integrated to double
   precision (the reference's QAG-41 stops at relative 1e-5); NaN for an
   invalid model. */
int pofd2d_get_bonus(pofd_ctx *ctx, uint32_t nwalk, double *out);
/* Priors and regularisation of calcLikeDouble::getLogLike
   (src/calcLikeDouble.cc:1230-1280), added on the device to every 2D batch
   after the dataset sum: flags POFD_PRIOR2_SIGMA1/2 (widths sigma_width[2],
   mean 1), POFD_PRIOR2_CFIRB1/2 (cfirb[4] = mean1, sigma1, mean2, sigma2,
   on the mean flux per area of each band), POFD_PRIOR2_POISSON1/2
   (poisson[4], on the mean flux^2 per area), and the band-1 knot
   regularisation when reg_alpha > 0. */
enum {
  POFD_PRIOR2_SIGMA1 = 1,
  POFD_PRIOR2_SIGMA2 = 2,
  POFD_PRIOR2_CFIRB1 = 4,
  POFD_PRIOR2_CFIRB2 = 8,
  POFD_PRIOR2_POISSON1 = 16,
  POFD_PRIOR2_POISSON2 = 32
};
int pofd2d_set_priors(pofd_ctx *ctx, int flags, const double *sigma_width,
                      const double *cfirb, const double *poisson,
                      double reg_alpha);

/* ---- parity and debugging (PDFactoryDouble::getPD output) ---- */
/* info[32]: status, split1, split2, sg1pos, sg2pos, sigma1, sigma2,
   integral, 1/integral, minflux1, minflux2, dataset logL, mn1, mn2, var1,
   var2, r0, K1, K2 (spectral bounds kept), split status of band 1 and of
   band 2 */
int pofd2d_get_debug(pofd_ctx *ctx, uint32_t walker, int beamset_id,
                     int field_id, double *info);
/* pd_out[n*n]: rotated, normalised P(D) (log2 when is_log) of the last
   batch; info[8] = minflux1, minflux2, dflux1, dflux2, split1, split2,
   sg1pos, sg2pos */
int pofd2d_get_pd(pofd_ctx *ctx, uint32_t walker, int beamset_id,
                  int field_id, int is_log, double *pd_out, double *info);

/* ======================================================================
   Device-resident ensemble sampler: the affine-invariant stretch move of
   affineEnsemble (src/affineEnsemble.cc:116-126, 1144-1420) with the
   walkers, proposals, acceptance and chain kept on the GPU; the likelihood
   of each half-ensemble is one batched launch of the 1D or 2D path.
   Differences from the reference: random numbers come from Philox4x32-10
   (counter-based, reproducible per seed) instead of ran (src/ran.cc); a
   proposal that stays outside the limits after 40 redraws, or whose
   likelihood fails (split, invalid model, non-finite mean or value), is
   rejected and counted per walker where the reference aborts the run.
   ====================================================================== */
typedef struct pofd_ens pofd_ens;
enum {
  POFD_TARGET_1D = 1,     /* pofd1d_loglike_batch_device on the context */
  POFD_TARGET_2D = 2,     /* pofd2d_loglike_batch_device on the context */
  POFD_TARGET_GAUSS = 3,  /* diagonal Gaussian (validation target) */
  POFD_TARGET_HOST = 4    /* the caller's function (pofd_ens_set_host_target) */
};
/* nwalkers even; npar parameters per walker (<= 64, bonus ones included,
   the theta stride of the likelihood); scalefac the stretch scale a > 1
   (default 2 in the reference, src/pofd_affine_mcmc.cc:39) */
int pofd_ens_create(pofd_ens **out, pofd_ctx *ctx, int target,
                    uint32_t nwalkers, uint32_t npar, double scalefac,
                    uint64_t seed);
void pofd_ens_destroy(pofd_ens *ens);
const char *pofd_ens_last_error(const pofd_ens *ens);
/* A likelihood the caller evaluates (POFD_TARGET_HOST): replaces the
   virtual affineEnsemble::getLogLike(const paramSet&, bool&) and
   fillBonusParams(paramSet&, bool) of a subclass (include/affineEnsemble.h:
   175,179) for likelihoods that are not on the device.  For every
   half-ensemble step the sampler copies the n proposals (rows of stride
   floats) to the host and calls fn once; fn writes loglike[n], status[n]
   (0 accepted as a value, nonzero: the proposal is rejected, as a
   rejected or invalid parameter set is) and, with nbonus > 0, bonus[n *
   nbonus]: the values of the bonus parameters (in parameter order) that an
   accepted step stores.  A nonzero return aborts the run with that code.
   fn runs on the calling thread inside pofd_ens_run / set_positions. */
typedef int (*pofd_host_loglike_fn)(void *user, const float *theta,
                                    uint32_t n, uint32_t stride,
                                    double *loglike, int32_t *status,
                                    double *bonus);
int pofd_ens_set_host_target(pofd_ens *ens, pofd_host_loglike_fn fn,
                             void *user, uint32_t nbonus);
/* state[npar]: 0 fit, 1 fixed, 2 bonus (mcmc_affine::FIXED / BONUS) */
int pofd_ens_set_param_state(pofd_ens *ens, const uint8_t *state);
/* The acceptance exponent of z: fit_dimension 0 (default) uses the
   reference's nparams - nfixed - 1, which counts bonus parameters
   (src/affineEnsemble.cc:1398); 1 uses the number of fit parameters less
   one, which the stretch move's detailed balance requires when bonus
   parameters are present.  *exponent (optional) receives the value. */
int pofd_ens_set_exponent_mode(pofd_ens *ens, int fit_dimension,
                               float *exponent);
/* parameter limits (NaN or NULL: none), areParamsValid */
int pofd_ens_set_limits(pofd_ens *ens, const float *lower, const float *upper);
int pofd_ens_set_gauss_target(pofd_ens *ens, const double *mean,
                              const double *sigma);
/* initial positions theta[nwalkers * npar]; evaluates their likelihoods */
int pofd_ens_set_positions(pofd_ens *ens, const float *theta);
/* nsteps stretch-move steps (both halves each, doMasterStep) at the given
   temperature, storing every thin-th step (0: none) */
int pofd_ens_run(pofd_ens *ens, uint32_t nsteps, double temperature,
                 uint32_t thin);
int pofd_ens_get_state(pofd_ens *ens, float *theta, double *loglike,
                       uint32_t *naccept, uint32_t *nerr);
int pofd_ens_chain_length(const pofd_ens *ens, uint32_t *nstored);
/* stored steps [first, first+count): theta[count*nwalkers*npar],
   loglike[count*nwalkers] */
int pofd_ens_get_chain(pofd_ens *ens, uint32_t first, uint32_t count,
                       float *theta, double *loglike);
int pofd_ens_clear_chain(pofd_ens *ens);
/* autocorrelation time of parameter `param` over the stored steps, from
   the walker-averaged series (affineChainSet::getAcor,
   src/affineChainSet.cc:959-984) */
int pofd_ens_acor(pofd_ens *ens, uint32_t param, double *tau, double *mean,
                  double *sigma);
/* n draws of the stretch factor z (affineEnsemble::generateZ) */
int pofd_ens_sample_z(pofd_ens *ens, uint32_t n, double *out);
/* Goodman's ACOR on a series (affineChainSet::acor,
   src/affineChainSet.cc:880-950): POFD_E_ARG if n < 50 */
int pofd_acor(const float *x, uint32_t n, double *mean, double *sigma,
              double *tau);

/* ======================================================================
   Delayed acceptance with a learned surrogate of the log likelihood.  No
   reference counterpart: the network is an MLP nin -> H -> H -> H -> 1 with
   GELU activations, bfloat16 weights and activations, float accumulation
   and float master weights, trained on the GPU with Adam (Kingma and Ba,
   2015) on (theta, exact logL) pairs, inputs and targets standardised.  In
   the sampler it is the first stage of a two-stage acceptance (Christen
   and Fox, 2005), so the chain keeps the exact posterior; the exact
   likelihood runs only for the proposals the first stage passes.
   ====================================================================== */
typedef struct pofd_mlp pofd_mlp;
/* nin <= 64 inputs, hidden <= 512 units */
int pofd_mlp_create(pofd_mlp **out, uint32_t nin, uint32_t hidden,
                    uint64_t seed);
void pofd_mlp_destroy(pofd_mlp *mlp);
const char *pofd_mlp_last_error(const pofd_mlp *mlp);
int pofd_mlp_info(const pofd_mlp *mlp, uint32_t *nin, uint32_t *hidden);
/* parameter vector: W1[H x nin], b1[H], W2[H x H], b2, W3[H x H], b3,
   W4[1 x H], b4[1] at offsets[8] (each segment padded to 8 floats) */
int pofd_mlp_nparams(const pofd_mlp *mlp, uint32_t *n);
int pofd_mlp_layout(const pofd_mlp *mlp, uint64_t *offsets);
int pofd_mlp_get_params(pofd_mlp *mlp, float *w);
int pofd_mlp_set_params(pofd_mlp *mlp, const float *w);
/* inputs x are whitened as z = W (x - xmean), W row-major nin x nin, before
   their bfloat16 rounding; the prediction is out*ystd + ymean */
int pofd_mlp_set_whitening(pofd_mlp *mlp, const float *xmean,
                           const float *wmat, double ymean, double ystd);
/* the diagonal case: z_i = (x_i - xmean_i) / xstd_i */
int pofd_mlp_set_standardization(pofd_mlp *mlp, const float *xmean,
                                 const float *xstd, double ymean,
                                 double ystd);
/* x[n x nin]; out[n] in log-likelihood units */
int pofd_mlp_predict(pofd_mlp *mlp, const float *x, uint32_t n, double *out);
/* the gradient of the mean squared error on standardised targets with
   respect to every parameter (grad[nparams]) and the loss, for one batch */
int pofd_mlp_loss_grad(pofd_mlp *mlp, const float *x, const double *y,
                       uint32_t n, float *grad, double *loss);
/* Adam training (step size decayed along a cosine to 5% of lr): the
   whitening from the mean and covariance of the training rows (W = L^-1,
   L L^T the covariance), targets standardised, the last holdout_frac of the
   rows (after one shuffle) held out; RMS errors of the trained network in
   log-likelihood units */
int pofd_mlp_train(pofd_mlp *mlp, const float *x, const double *y,
                   uint32_t n, uint32_t epochs, uint32_t batch, float lr,
                   float holdout_frac, uint64_t seed, double *train_rms,
                   double *holdout_rms);
/* attach a trained surrogate to the sampler: its inputs are the ensemble
   parameters param_index[nidx]; delayed acceptance is switched on only when
   holdout_rms <= gate (0.5 in log-likelihood units is the intended gate),
   *enabled tells which; mlp NULL switches it off */
int pofd_ens_set_surrogate(pofd_ens *ens, pofd_mlp *mlp,
                           const uint32_t *param_index, uint32_t nidx,
                           double gate, double holdout_rms, int *enabled);
/* Surrogate steps per half-step while a surrogate is on (no counterpart in
   the reference; default 1).  1 is the delayed-acceptance step above.  With
   nsteps > 1 each walker of a half first runs nsteps stretch moves against
   the other half accepted on the surrogate alone, then its end point is
   accepted with exp(((L(y) - L(x)) - (S(y) - S(x)))/T) on the exact log
   likelihood L (Liu's surrogate transition method): the chain still
   samples the exact posterior, every stored log likelihood is exact, one
   exact evaluation per walker whose inner chain moved, and a half-step
   moves each walker up to nsteps times.  nsteps in [1, 64]. */
int pofd_ens_set_surrogate_steps(pofd_ens *ens, uint32_t nsteps);
/* zero the per-walker acceptance and rejection counters (the end of a
   burn-in) */
int pofd_ens_reset_counts(pofd_ens *ens);
/* proposals made and exact likelihood evaluations since set_positions */
int pofd_ens_counts(pofd_ens *ens, uint64_t *n_proposed, uint64_t *n_exact);

#ifdef __cplusplus
}
#endif
#endif
