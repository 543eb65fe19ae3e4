// Declarations shared by the host code and the HIP kernels of the 2D
// (two-band) path.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>

#include <hip/hip_runtime.h>

#include "../../include/pofd_affine_amd.h"

namespace pofd {
// Gauss-Legendre nodes and weights on [-1, 1] (capi.cpp)
void gauss_legendre(int n, double* x, double* w);
}  // namespace pofd

namespace pofd2 {

constexpr int kMaxK = 32;          // knots, sigma knots, offset knots (each)
constexpr int kMaxFields2 = 8;     // datasets per beam set
constexpr int kLikeBlocks = 64;    // workgroups per (walker, dataset) in like2d
constexpr int kRowBlocks2 = 64;    // workgroups per walker in the inverse row transforms
// per-walker model block: header, knot spline, sigma and offset interpolants
// (4 coefficients per interval each), raw sigma and offset values
constexpr int kModelStride = 8 + 14 * kMaxK;
constexpr int kStats = 16;
// dynamic LDS ceiling of the R-grid kernel (3 doubles per beam element)
constexpr size_t kRgridMaxLds = 144 * 1024;
constexpr int kDbg2 = 32;

enum {  // stats slots (per walker)
  ST2_MN1 = 0, ST2_MN2, ST2_VAR1, ST2_VAR2, ST2_SHIFT1, ST2_SHIFT2,
  ST2_DOSHIFT1, ST2_DOSHIFT2, ST2_R0, ST2_K1, ST2_K2
};
enum {  // debug slots (per walker and dataset)
  D2_STATUS = 0, D2_SPLIT1, D2_SPLIT2, D2_SG1POS, D2_SG2POS, D2_SIGMA1,
  D2_SIGMA2, D2_INTEGRAL, D2_ITOT, D2_MINFLUX1, D2_MINFLUX2, D2_LL,
  D2_MN1, D2_MN2, D2_VAR1, D2_VAR2, D2_R0, D2_K1, D2_K2,
  D2_ST1, D2_ST2  // split status of each axis
};

// one sign component (pp, pn, np, nn) of the beam, device view
struct BeamComp {
  int n;                    // elements (histogram bins or pixels), 0 if absent
  const double* ieta1;      // 1/beam1 per element
  const double* w;          // histogram weights (nullptr for raw pixels)
  const double* logratio;   // log(beam1/beam2) per element
  double maxf1;             // maxknot * max(beam1): R is zero beyond
};

struct Field2Dev {
  double sigma_base1, sigma_base2, like_offset, like_norm;
  int binned;
  const uint32_t* bins;     // nb1 * nb2
  uint32_t nb1, nb2;
  double bincent01, bindelta1, bincent02, bindelta2;
  const double* d1;         // npix pixel pairs (unbinned)
  const double* d2;
  uint64_t npix;
};

// one flux grid for rgrid2d (theta-independent tables)
struct RGrid {
  int n1, n2;
  const double* f1;         // n1 band-1 fluxes
  const double* f2inv;      // n2: 1/|f2| or 0 (f2 == 0)
  const double* f2log;      // n2: log|f2|
  const int* f2sgn;         // n2: 0 for f2 > 0, 1 for f2 < 0
  double* out;              // + walker * wstride + i * ostride + j
  size_t wstride;
  int ostride;
  double scale;
};

struct Set2Dev {
  int n, nedge, edge_integ;
  int nk, ns, no;
  double logfac;                  // log2(10.0)
  double knots[kMaxK], logknots[kMaxK];
  double sigmaknots[kMaxK], offsetknots[kMaxK];
  double min_sigma;
  double maxsig1, maxsig2;
  double prefac;                  // (pixsize/3600)^2 / sqrt(2 pi)
  int hasneg1;
  BeamComp comp[4];
  double dflux1, dflux2, iflux1, iflux2;
  const double* rflux1;           // n
  const double* rflux2;           // n
  const double* edge1;            // nedge, log-spaced (use_edge_log_x)
  double iRxnorm, iRynorm, iR00norm;
  const double2* tw_half;         // n/2: e^{-2 pi i k/(n/2)}
  const double2* tw_n;            // n:   e^{-2 pi i k/n}
  double spectral_cut;
  int nfields;
  Field2Dev fields[kMaxFields2];
  // counts integrals: breakpoints (log2 of the band-1, sigma and offset
  // knots inside the band-1 range, sorted) and Gauss-Legendre rule
  int bonus;                      // this launch also integrates the counts
  int nbp;
  double bpts[3 * kMaxK];
  double glx[32], glw[32];
};

struct Work2 {
  int nwalk;
  double* model;                  // W * kModelStride
  double* stats;                  // W * kStats
  double* R;                      // W * n * n
  double* e00;                    // W * nedge * nedge
  double* ex;                     // W * nedge * n
  double* ey;                     // W * n * nedge
  double* rx;                     // W * edge_stride (n)
  double* ry;
  double* r00rows;
  double2* Y;                     // W * n * ystride   rows' spectra
  double2* YT;                    // W * (n/2+1) * n   transposed
  double2* PR;                    // W * (n/2+1) * n   exp(R^ - R^0) [k2][k1]
  double2* Q;                     // W * (n/2+1) * n   per dataset, inverse columns
  double2* QT;                    // W * n * ystride
  double* P;                      // W * nf * n * n    P(D) per dataset
  double* rsum1;                  // W * nf * edge_stride
  double* rsum2;
  double* rowA;
  double* rowB;
  double* colpart;                // W * kRowBlocks2 * n  column partial sums
  double* rstat_part;             // W * nparts * 4
  double* lpart;                  // W * nf * kLikeBlocks
  double* dbg;                    // W * nf * kDbg2
  double* bonus;                  // W * 8: NS, S1, S2, S1^2, S2^2 per area
  double* rpieces;                // W * rpieces_stride: per row of the R grid
                                  //   being filled, rpieces_w values (3 per
                                  //   beam element: weight, centre, width)
  int rpieces_w;
  size_t R_stride, ex_stride, edge_stride, Y_stride, YT_stride, PR_stride,
      Q_stride, QT_stride, P_stride, colpart_stride, rpieces_stride;
  int ystride;                    // n/2 + 1 (complex)
};

// two-band state hung off pofd_ctx (capi2d.cpp) and the context accessors
// that capi.cpp provides
struct Ctx2;
void ctx2_free(Ctx2* c2);
}  // namespace pofd2
struct pofd_ctx;
namespace pofd2 {
Ctx2*& ctx2_of(pofd_ctx* c);
int ctx_device(const pofd_ctx* c);
double ctx_spectral_cut(const pofd_ctx* c);
int ctx_fail(pofd_ctx* c, int code, const std::string& msg);
int ctx_stage_events(pofd_ctx* c, hipEvent_t** ev);

// launcher (pofd2d_kernel.hip)
int launch_loglike2d(const Set2Dev& s, const Work2& wk, const RGrid& gmain,
                     const RGrid& g00, const RGrid& gx, const RGrid& gy,
                     const float* d_theta, uint32_t stride, double* d_ll,
                     int32_t* d_st, int accumulate, void* stream,
                     hipEvent_t* ev);

// priors and regularisation of calcLikeDouble::getLogLike
// (src/calcLikeDouble.cc:1230-1280): flags kP2Sigma1/2, kP2Cfirb1/2,
// kP2Poisson1/2; regularisation of the band-1 knots when reg_alpha > 0
enum { kP2Sigma1 = 1, kP2Sigma2 = 2, kP2Cfirb1 = 4, kP2Cfirb2 = 8, kP2Poisson1 = 16,
       kP2Poisson2 = 32 };
struct Prior2Dev {
  int flags;
  double sigma_width[2], cfirb_mean[2], cfirb_sigma[2], poisson_mean[2], poisson_sigma[2];
  double reg_alpha;
};
// after every beam set of a batch: one thread per walker; s0/w0 are beam
// set 0's (the counts integrals)
int launch_prior2d(const Set2Dev& s0, const Work2& w0, const Prior2Dev& pr, const float* d_theta,
                   uint32_t stride, uint32_t nwalk, double* d_ll, const int32_t* d_st,
                   void* stream);

}  // namespace pofd2
