// Declarations shared by the ensemble sampler's host code and kernels.
#pragma once

#include <cstddef>
#include <cstdint>

#include <hip/hip_runtime.h>

#include "mlp_internal.h"

namespace pofd_ens_dev {

constexpr int kMaxPar = 64;
enum { kStateFit = 0, kStateFixed = 1, kStateBonus = 2 };
constexpr int kBadLike = 4;  // POFD_E_BADLIKE: a likelihood value, not an error

struct EnsDev {
  int nwalk, npar;
  float dexp;                    // nparams - nfixed - 1 (affineEnsemble.cc:1398)
  double scalefac;
  uint64_t seed;
  uint8_t state[kMaxPar];
  uint8_t has_lower[kMaxPar], has_upper[kMaxPar];
  float lower[kMaxPar], upper[kMaxPar];
  double gmean[kMaxPar], gisig[kMaxPar];  // Gaussian test target
  float* pos;        // nwalk x npar   current positions
  double* ll;        // nwalk          current log likelihoods
  float* prop;       // nwalk/2 x npar proposals of one half
  double* llp;       // nwalk/2
  int32_t* stp;      // nwalk/2        likelihood status of the proposals
  float* z;          // nwalk/2
  int32_t* pvalid;   // nwalk/2        proposal inside the limits
  uint32_t* naccept; // nwalk
  uint32_t* nerr;    // nwalk          rejected for an invalid proposal or likelihood
  float* chain;      // capacity x nwalk x npar
  double* chainll;   // capacity x nwalk
  // delayed acceptance (surrogate first stage)
  double* llsur;     // nwalk          surrogate log likelihood of the walkers
  double* llsp;      // nwalk/2        surrogate of the proposals
  int32_t* pass;     // nwalk/2        compact index of a survivor, -1 if rejected
  int32_t* cidx;     // nwalk/2        compact list of the passing proposals
  int32_t* npass;    // 1
  float* cprop;      // nwalk/2 x npar compact proposals
  double* cll;       // nwalk/2        their exact log likelihoods
  int32_t* cst;      // nwalk/2
  unsigned long long* nexact;  // 1: exact evaluations requested by stage 1
  // surrogate transition (an inner chain of several surrogate steps)
  float* yin;        // nwalk/2 x npar inner-chain positions
  double* lsy;       // nwalk/2        their surrogate log likelihoods
  int32_t* moved;    // nwalk/2        the inner chain has moved
  // counts integrals of the last likelihood batch, bonus_stride per batch
  // row, [0] the source count: the first bonus_count bonus parameters of
  // an accepted step take entries 1.. (1D: mean flux and flux^2 per area,
  // pofdMCMC::fillBonusParams; 2D: both bands' flux and flux^2 per area,
  // pofdMCMCDouble::fillBonusParams); null: NaN
  const double* bonus_src;
  int bonus_stride, bonus_count;
};

// from: the walkers' positions (null: the current ones); draw0: first draw
int launch_propose(const EnsDev& e, uint32_t half, uint32_t step, hipStream_t st,
                   const float* from = nullptr, uint32_t draw0 = 0);
int launch_gauss(const EnsDev& e, const float* x, int n, double* ll, int32_t* stv,
                 hipStream_t st);
int launch_accept(const EnsDev& e, uint32_t half, uint32_t step, double inv_temp, int store,
                  uint32_t slot, hipStream_t st);
int launch_zsample(const EnsDev& e, uint32_t n, double* out, hipStream_t st);
// delayed acceptance: first stage on the surrogate, compaction of the
// passing proposals, second stage on their exact likelihoods, chain store
// (ys: the candidate rows, the proposals when null; the survivors' rows
// are gathered into cprop by the first stage, and the second stage stores
// the half's walkers into chain slot `slot` when `store`)
int launch_stage1(const EnsDev& e, uint32_t half, uint32_t step, double inv_temp,
                  hipStream_t st, int moved_only = 0, const float* ys = nullptr);
int launch_stage2(const EnsDev& e, uint32_t half, uint32_t step, double inv_temp, int store,
                  uint32_t slot, hipStream_t st, const float* ys = nullptr,
                  const double* lsn = nullptr, uint32_t draw = 102);
// one inner step of the surrogate transition in one launch: the proposals
// from the inner positions `from` (draws from draw0), their surrogate
// predictions and their inner acceptance (uniform draw adraw)
int launch_inner_step(const EnsDev& e, uint32_t half, uint32_t step, const float* from,
                      uint32_t draw0, const pofd_mlp_dev::InferArgs& a, double inv_temp,
                      uint32_t adraw, hipStream_t st);
// surrogate transition: the inner positions from the walkers
int launch_st_init(const EnsDev& e, uint32_t half, hipStream_t st);
// bonus parameters of all walkers from the batch that evaluated them in order
int launch_bonus_init(const EnsDev& e, hipStream_t st);

}  // namespace pofd_ens_dev
