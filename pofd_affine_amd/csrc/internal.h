// Internal declarations shared by the host code and the HIP kernels.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/pofd_affine_amd.h"

namespace pofd {

// beam after beam::readFile (+ makeHistogram), include/beam.h:28-64
struct Beam {
  double pixsize = 0.0, minval = 0.0;
  std::vector<double> pos, neg, posinv, neginv;  // sorted |pixel| values
  double totpos = 0.0, totpossq = 0.0, totneg = 0.0, totnegsq = 0.0;
  uint32_t nbins = 0;
  bool is_pos_hist = false, is_neg_hist = false;
  std::vector<double> posw, posv, negw, negv;    // hist weights, inverse vals
  int set(const double* pix, size_t n, double pixsize, double minval,
          std::string& err);
  int make_histogram(uint32_t nbins, std::string& err);
};

void data_minmax(const double* d, size_t n, double& mn, double& mx);
// Gauss-Legendre nodes and weights on [-1, 1]
void gauss_legendre(int n, double* x, double* w);

// Limits of the 1D kernels.
constexpr int kThreads = 1024;   // threads per workgroup (one walker)
constexpr int kM1 = 256;         // column transform length
constexpr int kChunk = 16;       // columns per LDS chunk
constexpr int kRowsPerPass = 8;  // spectral rows accumulated per forward pass
constexpr int kMaxKnots = 64;
constexpr int kBlk = 32;         // P(D) values per min/sum block
constexpr int kWarm = 64;        // most warm-up steps of the truncated recurrences
constexpr int kDbg = 32;         // debug scalars per (walker, field)
constexpr int kState = 16;       // per-walker scalars handed between kernels
constexpr int kPixBlocks = 64;   // workgroups per (walker, field), unbinned
constexpr int kGL = 32;          // Gauss-Legendre points per knot interval
constexpr int kIPart = 8;        // partial statistics per inverse part
constexpr int kMaxSlots = 256;   // (walker, part) slots of the split transforms
// device scratch of the allocation-free batch entry point, per GPU
constexpr size_t kNoallocScratchBytes = size_t(512) << 20;

// One field (dataset) of a beam set, device view.
struct FieldDev {
  double sigma_base;
  double like_offset;
  double like_norm;
  int binned;
  const uint32_t* bins;   // nbins
  uint32_t nbins;
  double bincent0, bindelta;
  const double* pix;      // npix (sorted ascending), unbinned only
  uint64_t npix;
};

constexpr int kMaxFields = 16;

// Everything the kernels need for one beam set; theta-independent.
struct BeamsetDev {
  int N;        // nflux
  int M;        // N/2
  int M2;       // M / 256 (number of columns)
  int ninterp;
  int nknots;
  int nbeam;    // beam entries (histogram bins or raw pixels), positive
  int nbeamn;   // beam entries of the negative beam (after the positive)
  int nseg;     // node segments: 1, or 2 with the negative beam
  int haspos;   // segment 0 holds a positive beam
  int minit, maxit;       // positive fill range of R indices (PDFactory.cc:569-574)
  int minitn, maxitn;     // negative (wrapped) fill range (:586-613)
  int wrapRidx;           // last index of a non-negative flux
  double dflux;
  double minflux_R;
  double mval;            // flux of index i > wrapRidx is i*dflux + mval
  double logfac;          // log2(10.0) as glibc computes it
  double prefac;          // (pixsize/3600)^2
  double maxsigma_base;
  double iflux;           // two_pi / (N * dflux)
  double spectral_cut;    // drop components with sigma^2 w^2/2 > cut
  double logknots[kMaxKnots];
  int bonus;              // this launch also integrates the counts
  double glx[kGL], glw[kGL];  // Gauss-Legendre nodes and weights on [-1, 1]
  // node tables, per segment (positive beam, then negative beam):
  const double* xnode;    // nseg*ninterp interpolation fluxes
  const double* ginv;     // nseg*(ninterp-1): 1/h_i  (GSL cspline_init g_i)
  const double* lalpha;   // nseg*(ninterp-2): LDL^T diagonal of the node system
  const double* lgamma;   // nseg*(ninterp-2): LDL^T sub-diagonal multipliers
  int kwarm[2];           // per segment: warm-up steps of the node system's
                          // truncated recurrences, kwarm >= 64 / -log2(max
                          // multiplier) (at most 64: the multipliers are < 1/2)
  const double* kfac;     // 4 x kMaxKnots, the knot spline's fixed factors:
                          // h_i, 1/h_i, LDL^T diagonal, LDL^T multipliers
  const int8_t* rk;       // (nbeam+nbeamn)*ninterp: knot interval or -1 (skip)
  const double* rdelx;    // (nbeam+nbeamn)*ninterp: log2|S| - logknots[k]
  const double* bw;       // nbeam+nbeamn weights (1.0 for a raw beam)
  const double* bieta;    // nbeam+nbeamn inverse |beam| values
  const uint16_t* kst;    // N: segment*ninterp + node interval of every grid flux
  const double2* W256;    // 256: exp(-2 pi i s/256)
  const double2* Wlo;     // 256: exp(-2 pi i s/M)
  const double2* Wm2;     // M2: exp(-2 pi i s/M2)
  const double2* em1;     // N/2+1: exp(-2 pi i t/N) - 1, without cancellation
                          // (headn points of the head, headn <= M2)
  // grid points [0, headn) -- and [N - headn, N) when headneg -- hold R's
  // largest values (fluxes near zero); they are left out of the transform
  // and their part of R^(k) - R^(0) is summed directly (forward_tail)
  int headn, headneg;
  int nfields;
  FieldDev fields[kMaxFields];
};

// Per-walker device workspace layout (doubles), sized by the host.
struct WorkDev {
  double* coef;      // W * nseg*ninterp*4 : y, b, c, d per interval and segment
  double* ynode;     // W * nseg*ninterp   : log2 R at the nodes
  double2* zrows;    // W * (1+nfields) * M: forward rows, then the noisy
                     //   rows of each dataset (M complex each)
  double2* prtrans;  // W * (M+1)          : exp(R^ - R^0) with phase
  double* pofd;      // W * nfields * N    : clamped, unrotated P(D)
  double* bmin;      // W * nfields * N/32
  double* bsum;      // W * nfields * N/32
  double* dbg;       // W * nfields * kDbg : debug scalars
  double* state;     // W * kState         : per-walker scalars
  double* upart;     // W * nfields * kPixBlocks : unbinned partial sums
  double* bonus;     // W * 4: int n dS, int S n dS, int S^2 n dS
  // walkers split over several workgroups (batches of at most 128): one
  // slot per (walker, part), nslots = min(256, 4 W), walkers x parts <= 256
  double2* zpart;    // nslots * M : partial kept rows of the forward transform
  double* mpart;     // nslots * 2 : partial moments (sum S R, sum S^2 R)
  double* ipart;     // nslots * nfields * kIPart : partial inverse statistics
  uint32_t nslots;
  size_t coef_stride, ynode_stride, zrows_stride, spec_stride, pofd_stride,
      blk_stride, zpart_stride;
  // the batch of one launch: n walkers, read from dcount on the device when
  // it is set (the grids then cover the largest possible batch and the
  // workgroups past it return), else nhost; force_parts as pofd1d_set_parts
  const int32_t* dcount;
  int nhost;
  int force_parts;
  // measured workgroups per walker by batch size (kMaxSlots + 1 entries,
  // 0 = the size rule), or null before pofd1d_autotune_parts
  const uint8_t* parts_tab;
};

// The transforms' column chunks fall in kGroups fixed groups, each summed
// in chunk order, and the group sums combine as (G0 + G1) + (G2 + G3).  A
// split transform's part covers whole groups (1, 2 or 4 parts), so a
// walker's sums, and with them its log likelihood, are the same bits for
// every batch size.
constexpr int kGroups = 4;

// workgroups per walker of the split transforms for a batch of n walkers
// (the same rule on the host and in the kernels): the forced count, else
// the measured table's entry, else min(4, 256 / n); 3 becomes 2, and the
// count halves while n x parts exceeds the workspace's slots
__host__ __device__ inline int batch_parts(int n, int force, uint32_t nslots,
                                           const uint8_t* tab) {
  int parts = force;
  if (parts == 0 && tab && n > 0 && n <= kMaxSlots) parts = tab[n];
  if (parts == 0) {
    parts = (n > 0) ? 256 / n : 4;
    parts = parts < 1 ? 1 : (parts > 4 ? 4 : parts);
  }
  if (parts == 3) parts = 2;
  while (parts > 1 && static_cast<uint32_t>(n * parts) > nslots) parts >>= 1;
  return parts;
}

}  // namespace pofd

struct pofd_beam {
  pofd::Beam beam;
};

// launcher implemented in pofd1d_kernel.hip
namespace pofd {
// ev: nullptr, or POFD_NSTAGES+1 events recorded around the stages
// nwalk: the batch, or with d_count its largest possible size
int launch_loglike1d(const BeamsetDev& bs, const WorkDev& wk, int beamset_idx,
                     const float* d_theta, uint32_t nwalk, const int32_t* d_count,
                     uint32_t stride, double* d_loglike, int32_t* d_status,
                     int accumulate, int force_parts, const uint8_t* parts_tab_host,
                     const uint8_t* parts_tab_dev, void* stream, hipEvent_t* ev);
// priors and regularisation added to the summed log likelihood
// (calcLike::getLogLike, src/calcLike.cc:923-952); flags: kPriorSigma,
// kPriorCfirb, kPriorPoisson; regularisation when reg_alpha > 0
enum { kPriorSigma = 1, kPriorCfirb = 2, kPriorPoisson = 4 };
struct PriorDev {
  int flags;
  double sigma_width, cfirb_mean, cfirb_sigma, poisson_mean, poisson_sigma, reg_alpha;
};
// after every beam set of a batch: one thread per walker; bs0/wk0 are
// beam set 0's (the counts integrals), the batch size as for the batch
int launch_prior1d(const BeamsetDev& bs0, const WorkDev& wk0, const PriorDev& pr,
                   const float* d_theta, uint32_t stride, uint32_t nwalk, const int32_t* d_count,
                   double* d_loglike, const int32_t* d_status, void* stream);
// PD::getLogLike / PDDouble::getLogLike over a P(D) in device memory
// (pdlike_kernel.hip); d_scratch holds pd_like_scratch_doubles() doubles,
// the result lands in its last one
size_t pd_like_scratch_doubles();
int launch_pd_loglike1d(const double* d_pd, int n, double minflux, double dflux, int islog,
                        const double* d_pix, uint64_t npix, const uint32_t* d_bins,
                        uint32_t nbins, double bincent0, double bindelta, double* d_scratch,
                        void* stream);
int launch_pd_loglike2d(const double* d_pd, int n1, int n2, double minflux1, double dflux1,
                        double minflux2, double dflux2, int islog, const double* d_pix1,
                        const double* d_pix2, uint64_t npix, const uint32_t* d_bins,
                        uint32_t nb1, uint32_t nb2, double bincent01, double bindelta1,
                        double bincent02, double bindelta2, double* d_scratch, void* stream);
// the allocation-free entry point's scratch (noalloc.hip): resolve its
// address on `device` (current) once, then look it up (nullptr: none)
int noalloc_init(int device);
void* noalloc_scratch(int device);
int launch_materialize_pd(const double* d_pofd, int N, int split, double itot,
                          int is_log, int already_log, double* d_out,
                          void* stream);
}  // namespace pofd
