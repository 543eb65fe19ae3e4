// Host side of the device-resident ensemble sampler (C ABI pofd_ens_*) and
// the autocorrelation estimate of affineChainSet (pofd_acor).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pofd_affine_amd.h"
#include "ensemble_internal.h"

using pofd_ens_dev::EnsDev;

// counts integrals of the last batch on the device (capi.cpp, capi2d.cpp)
const double* pofd1d_bonus_device(pofd_ctx* c);
extern "C" const double* pofd2d_bonus_device(pofd_ctx* c);

// surrogate predictions on the device (mlp.cpp)
int pofd_mlp_predict_device(pofd_mlp* s, const float* src, int stride, const int* idx, int rows,
                            double* out, hipStream_t st);

struct pofd_ens {
  pofd_ctx* ctx = nullptr;
  int target = 0;
  uint32_t nwalk = 0, npar = 0;
  EnsDev e{};
  hipStream_t stream = nullptr;
  void* dbuf[11] = {};
  uint32_t capacity = 0, nstored = 0, step = 0;
  bool positioned = false;
  bool fit_dimension = false;  // exponent from the fit parameters only
  // delayed acceptance
  pofd_mlp* mlp = nullptr;
  bool use_da = false;
  int* sidx = nullptr;         // parameter index of each surrogate input
  void* dabuf[9] = {};
  // surrogate steps per half-step: 1 is delayed acceptance, more is the
  // surrogate transition's inner chain
  uint32_t st_steps = 1;
  void* stbuf[3] = {};
  uint64_t n_exact = 0, n_proposed = 0;
  // host likelihood (POFD_TARGET_HOST): the caller's function, host staging
  // and the device copy of the bonus values it returns
  pofd_host_loglike_fn host_fn = nullptr;
  void* host_user = nullptr;
  uint32_t host_nbonus = 0;
  std::vector<float> host_theta;
  std::vector<double> host_ll, host_bonus;
  std::vector<int32_t> host_st;
  double* host_bonus_dev = nullptr;
  std::string err;
  ~pofd_ens() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (void* p : dbuf)
      if (p) (void)hipFree(p);
    if (e.chain) (void)hipFree(e.chain);
    if (e.chainll) (void)hipFree(e.chainll);
    for (void* p : dabuf)
      if (p) (void)hipFree(p);
    for (void* p : stbuf)
      if (p) (void)hipFree(p);
    if (sidx) (void)hipFree(sidx);
    if (host_bonus_dev) (void)hipFree(host_bonus_dev);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// random-draw numbers of the surrogate transition, apart from those of the
// plain and delayed-acceptance steps (0-41, 100-102): inner step i
// proposes with draws kInnerDraw0 + kInnerDrawStride i + (0 .. 41) and
// accepts with + kInnerAcceptDraw; the exact correction uses kCorrectDraw
constexpr uint32_t kInnerDraw0 = 1000, kInnerDrawStride = 64, kInnerAcceptDraw = 50;
constexpr uint32_t kCorrectDraw = 999;
constexpr uint32_t kMaxInnerSteps = 64;

int efail(pofd_ens* s, int code, const std::string& msg) {
  if (s) s->err = msg;
  return code;
}

int ehip(pofd_ens* s, hipError_t e, const char* what) {
  return efail(s, POFD_E_INTERNAL,
               std::string("pofd::hip: ") + what + ": " + hipGetErrorString(e));
}

void update_dexp(pofd_ens* s) {
  int nfixed = 0, nbonus = 0;
  for (uint32_t i = 0; i < s->npar; ++i) {
    if (s->e.state[i] == pofd_ens_dev::kStateFixed) ++nfixed;
    if (s->e.state[i] == pofd_ens_dev::kStateBonus) ++nbonus;
  }
  // affineEnsemble::emptyMasterQueue, src/affineEnsemble.cc:1398: the
  // reference's exponent counts all parameters, bonus ones included, less
  // the fixed; the stretch move's detailed balance needs the number of fit
  // parameters less one (the two agree when there are no bonus parameters)
  const int d = static_cast<int>(s->npar) - nfixed - (s->fit_dimension ? nbonus : 0);
  s->e.dexp = static_cast<float>(d - 1);
}

// the caller's likelihood on the host: the rows come back from the device,
// the function fills log likelihoods, statuses and bonus values, and they go
// to the device (the bonus values where the accept kernel reads them)
int evaluate_host(pofd_ens* s, const float* x, uint32_t n, double* ll, int32_t* st) {
  if (!s->host_fn) return efail(s, POFD_E_ARG, "pofd_ens_run: no host likelihood set");
  const size_t P = s->npar, nb = s->host_nbonus;
  s->host_theta.resize(n * P);
  s->host_ll.assign(n, 0.0);
  s->host_st.assign(n, 0);
  s->host_bonus.assign(n * (nb + 1), 0.0);
  hipError_t he = hipMemcpyAsync(s->host_theta.data(), x, n * P * sizeof(float),
                                 hipMemcpyDeviceToHost, s->stream);
  if (he == hipSuccess) he = hipStreamSynchronize(s->stream);
  if (he != hipSuccess) return ehip(s, he, "hipMemcpy(host proposals)");
  std::vector<double> bon(n * nb, 0.0);
  const int rc = s->host_fn(s->host_user, s->host_theta.data(), n, static_cast<uint32_t>(P),
                            s->host_ll.data(), s->host_st.data(), nb ? bon.data() : nullptr);
  if (rc) return efail(s, rc, "pofd_ens_run: the host likelihood failed");
  for (uint32_t r = 0; r < n; ++r)
    for (size_t b = 0; b < nb; ++b) s->host_bonus[r * (nb + 1) + 1 + b] = bon[r * nb + b];
  he = hipMemcpyAsync(ll, s->host_ll.data(), n * sizeof(double), hipMemcpyHostToDevice, s->stream);
  if (he == hipSuccess)
    he = hipMemcpyAsync(st, s->host_st.data(), n * sizeof(int32_t), hipMemcpyHostToDevice,
                        s->stream);
  if (he == hipSuccess && nb)
    he = hipMemcpyAsync(s->host_bonus_dev, s->host_bonus.data(), n * (nb + 1) * sizeof(double),
                        hipMemcpyHostToDevice, s->stream);
  // the host vectors are reused by the next call: the copies must finish
  if (he == hipSuccess) he = hipStreamSynchronize(s->stream);
  if (he != hipSuccess) return ehip(s, he, "hipMemcpy(host likelihoods)");
  s->e.bonus_src = nb ? s->host_bonus_dev : nullptr;
  s->e.bonus_stride = static_cast<int>(nb + 1);
  s->e.bonus_count = static_cast<int>(nb);
  return POFD_OK;
}

// evaluate n parameter rows x (device) into ll / st (device) on the stream
int evaluate(pofd_ens* s, const float* x, uint32_t n, double* ll, int32_t* st) {
  switch (s->target) {
    case POFD_TARGET_1D: {
      const int rc = pofd1d_loglike_batch_device(s->ctx, x, n, s->npar, ll, st, s->stream);
      s->e.bonus_src = pofd1d_bonus_device(s->ctx);  // the workspace may have moved
      s->e.bonus_stride = 4;
      s->e.bonus_count = 2;
      return rc;
    }
    case POFD_TARGET_2D: {
      const int rc = pofd2d_loglike_batch_device(s->ctx, x, n, s->npar, ll, st, s->stream);
      s->e.bonus_src = pofd2d_bonus_device(s->ctx);
      s->e.bonus_stride = 8;
      s->e.bonus_count = 4;
      return rc;
    }
    case POFD_TARGET_GAUSS:
      return pofd_ens_dev::launch_gauss(s->e, x, static_cast<int>(n), ll, st, s->stream)
                 ? POFD_E_INTERNAL
                 : POFD_OK;
    case POFD_TARGET_HOST:
      return evaluate_host(s, x, n, ll, st);
    default:
      return POFD_E_ARG;
  }
}

// surrogate log likelihoods of all walkers' current positions
int refresh_surrogate(pofd_ens* s) {
  if (!s->use_da || !s->positioned) return POFD_OK;
  if (pofd_mlp_predict_device(s->mlp, s->e.pos, static_cast<int>(s->npar), s->sidx,
                              static_cast<int>(s->nwalk), s->e.llsur, s->stream))
    return efail(s, POFD_E_INTERNAL, "pofd_ens: surrogate prediction failed");
  return POFD_OK;
}

// the exact log likelihoods of the first stage's survivors (cprop, *npass
// rows) into cll / cst; the 1D batch reads the count on the device
int exact_survivors(pofd_ens* s) {
  const uint32_t hw = s->nwalk / 2;
  EnsDev& e = s->e;
  if (s->target == POFD_TARGET_1D) {
    // the survivors' count stays on the device: no host round trip
    int rc = pofd1d_loglike_batch_device_count(s->ctx, e.cprop, hw, e.npass, s->npar, e.cll,
                                               e.cst, s->stream);
    if (rc) return efail(s, rc, "pofd_ens_run: likelihood launch failed");
    e.bonus_src = pofd1d_bonus_device(s->ctx);
    e.bonus_stride = 4;
    e.bonus_count = 2;
    return POFD_OK;
  }
  int32_t npass = 0;
  hipError_t he = hipMemcpyAsync(&npass, e.npass, sizeof(int32_t), hipMemcpyDeviceToHost,
                                 s->stream);
  if (he == hipSuccess) he = hipStreamSynchronize(s->stream);
  if (he != hipSuccess) return ehip(s, he, "hipMemcpy(npass)");
  if (npass > 0) {
    int rc = evaluate(s, e.cprop, static_cast<uint32_t>(npass), e.cll, e.cst);
    if (rc) return efail(s, rc, "pofd_ens_run: likelihood launch failed");
  }
  return POFD_OK;
}

// one half-step with the delayed-acceptance first stage
int da_half(pofd_ens* s, uint32_t half, uint32_t step, double inv_temp, int store) {
  const uint32_t hw = s->nwalk / 2;
  if (pofd_ens_dev::launch_propose(s->e, half, step, s->stream))
    return efail(s, POFD_E_INTERNAL, "pofd_ens_run: propose launch failed");
  if (pofd_mlp_predict_device(s->mlp, s->e.prop, static_cast<int>(s->npar), s->sidx,
                              static_cast<int>(hw), s->e.llsp, s->stream))
    return efail(s, POFD_E_INTERNAL, "pofd_ens_run: surrogate prediction failed");
  if (pofd_ens_dev::launch_stage1(s->e, half, step, inv_temp, s->stream))
    return efail(s, POFD_E_INTERNAL, "pofd_ens_run: first stage launch failed");
  int rc = exact_survivors(s);
  if (rc) return rc;
  if (pofd_ens_dev::launch_stage2(s->e, half, step, inv_temp, store, s->nstored, s->stream))
    return efail(s, POFD_E_INTERNAL, "pofd_ens_run: second stage launch failed");
  return POFD_OK;
}

// one half-step of the surrogate transition: st_steps stretch moves per
// walker on the surrogate (the other half fixed), then one exact
// correction of the walkers whose inner chain moved
int st_half(pofd_ens* s, uint32_t half, uint32_t step, double inv_temp, int store) {
  EnsDev& e = s->e;
  if (pofd_ens_dev::launch_st_init(e, half, s->stream))
    return efail(s, POFD_E_INTERNAL, "pofd_ens_run: inner chain launch failed");
  pofd_mlp_dev::InferArgs ia;
  pofd_mlp_infer_args(s->mlp, s->sidx, &ia);
  for (uint32_t i = 0; i < s->st_steps; ++i) {
    const uint32_t d0 = kInnerDraw0 + kInnerDrawStride * i;
    // the inner step's proposals, predictions and acceptance in one launch
    if (pofd_ens_dev::launch_inner_step(e, half, step, e.yin, d0, ia, inv_temp,
                                        d0 + kInnerAcceptDraw, s->stream))
      return efail(s, POFD_E_INTERNAL, "pofd_ens_run: inner chain launch failed");
  }
  if (pofd_ens_dev::launch_stage1(e, half, step, inv_temp, s->stream, 1, e.yin))
    return efail(s, POFD_E_INTERNAL, "pofd_ens_run: compaction launch failed");
  int rc = exact_survivors(s);
  if (rc) return rc;
  if (pofd_ens_dev::launch_stage2(e, half, step, inv_temp, store, s->nstored, s->stream, e.yin,
                                  e.lsy, kCorrectDraw))
    return efail(s, POFD_E_INTERNAL, "pofd_ens_run: correction launch failed");
  return POFD_OK;
}

}  // namespace

extern "C" {

int pofd_ens_create(pofd_ens** out, pofd_ctx* ctx, int target, uint32_t nwalkers,
                    uint32_t npar, double scalefac, uint64_t seed) {
  if (!out) return POFD_E_ARG;
  *out = nullptr;
  if (nwalkers < 2 || (nwalkers & 1) || npar == 0 ||
      npar > static_cast<uint32_t>(pofd_ens_dev::kMaxPar) || !(scalefac > 1.0))
    return POFD_E_ARG;
  if ((target == POFD_TARGET_1D || target == POFD_TARGET_2D) && !ctx) return POFD_E_ARG;
  if (target != POFD_TARGET_1D && target != POFD_TARGET_2D && target != POFD_TARGET_GAUSS &&
      target != POFD_TARGET_HOST)
    return POFD_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return POFD_E_NOGPU;
  pofd_ens* s = new pofd_ens();
  s->ctx = ctx;
  s->target = target;
  s->nwalk = nwalkers;
  s->npar = npar;
  EnsDev& e = s->e;
  e.nwalk = static_cast<int>(nwalkers);
  e.npar = static_cast<int>(npar);
  e.scalefac = scalefac;
  e.seed = seed;
  for (int i = 0; i < pofd_ens_dev::kMaxPar; ++i) {
    e.gisig[i] = 1.0;
    e.gmean[i] = 0.0;
  }
  update_dexp(s);
  hipError_t he = hipStreamCreate(&s->stream);
  if (he != hipSuccess) {
    delete s;
    return POFD_E_INTERNAL;
  }
  const size_t W = nwalkers, H = nwalkers / 2, P = npar;
  const size_t sizes[11] = {W * P * sizeof(float), W * sizeof(double), H * P * sizeof(float),
                            H * sizeof(double),    H * sizeof(int32_t), H * sizeof(float),
                            H * sizeof(int32_t),   W * sizeof(uint32_t), W * sizeof(uint32_t),
                            W * sizeof(int32_t),   64};
  for (int i = 0; i < 11; ++i) {
    he = hipMalloc(&s->dbuf[i], sizes[i]);
    if (he == hipSuccess) he = hipMemset(s->dbuf[i], 0, sizes[i]);
    if (he != hipSuccess) {
      delete s;
      return POFD_E_INTERNAL;
    }
  }
  e.pos = static_cast<float*>(s->dbuf[0]);
  e.ll = static_cast<double*>(s->dbuf[1]);
  e.prop = static_cast<float*>(s->dbuf[2]);
  e.llp = static_cast<double*>(s->dbuf[3]);
  e.stp = static_cast<int32_t*>(s->dbuf[4]);
  e.z = static_cast<float*>(s->dbuf[5]);
  e.pvalid = static_cast<int32_t*>(s->dbuf[6]);
  e.naccept = static_cast<uint32_t*>(s->dbuf[7]);
  e.nerr = static_cast<uint32_t*>(s->dbuf[8]);
  *out = s;
  return POFD_OK;
}

void pofd_ens_destroy(pofd_ens* s) { delete s; }

const char* pofd_ens_last_error(const pofd_ens* s) { return s ? s->err.c_str() : ""; }

int pofd_ens_set_param_state(pofd_ens* s, const uint8_t* state) {
  if (!s || !state) return POFD_E_ARG;
  for (uint32_t i = 0; i < s->npar; ++i) {
    if (state[i] > pofd_ens_dev::kStateBonus)
      return efail(s, POFD_E_ARG, "pofd_ens_set_param_state: state must be 0, 1 or 2");
    s->e.state[i] = state[i];
  }
  update_dexp(s);
  return POFD_OK;
}

int pofd_ens_set_exponent_mode(pofd_ens* s, int fit_dimension, float* exponent) {
  if (!s) return POFD_E_ARG;
  s->fit_dimension = fit_dimension != 0;
  update_dexp(s);
  if (exponent) *exponent = s->e.dexp;
  return POFD_OK;
}

int pofd_ens_set_limits(pofd_ens* s, const float* lower, const float* upper) {
  if (!s) return POFD_E_ARG;
  for (uint32_t i = 0; i < s->npar; ++i) {
    s->e.has_lower[i] = (lower && !std::isnan(lower[i])) ? 1 : 0;
    s->e.lower[i] = s->e.has_lower[i] ? lower[i] : 0.0f;
    s->e.has_upper[i] = (upper && !std::isnan(upper[i])) ? 1 : 0;
    s->e.upper[i] = s->e.has_upper[i] ? upper[i] : 0.0f;
    if (s->e.has_lower[i] && s->e.has_upper[i] && s->e.lower[i] > s->e.upper[i])
      return efail(s, POFD_E_ARG, "pofd_ens_set_limits: lower limit above upper limit");
  }
  return POFD_OK;
}

int pofd_ens_set_gauss_target(pofd_ens* s, const double* mean, const double* sigma) {
  if (!s || !mean || !sigma) return POFD_E_ARG;
  for (uint32_t i = 0; i < s->npar; ++i) {
    if (!(sigma[i] > 0.0))
      return efail(s, POFD_E_ARG, "pofd_ens_set_gauss_target: sigma must be positive");
    s->e.gmean[i] = mean[i];
    s->e.gisig[i] = 1.0 / sigma[i];
  }
  return POFD_OK;
}

int pofd_ens_set_positions(pofd_ens* s, const float* theta) {
  if (!s || !theta) return POFD_E_ARG;
  const size_t W = s->nwalk, P = s->npar;
  hipError_t he = hipMemcpy(s->e.pos, theta, W * P * sizeof(float), hipMemcpyHostToDevice);
  if (he != hipSuccess) return ehip(s, he, "hipMemcpy(positions)");
  int32_t* st = static_cast<int32_t*>(s->dbuf[9]);
  int rc = evaluate(s, s->e.pos, s->nwalk, s->e.ll, st);
  if (rc) return efail(s, rc, "pofd_ens_set_positions: likelihood of the initial positions failed");
  std::vector<int32_t> hst(W);
  he = hipStreamSynchronize(s->stream);
  if (he == hipSuccess) he = hipMemcpy(hst.data(), st, W * sizeof(int32_t), hipMemcpyDeviceToHost);
  if (he != hipSuccess) return ehip(s, he, "hipMemcpy(status)");
  for (size_t w = 0; w < W; ++w)
    if (hst[w] != POFD_OK && hst[w] != POFD_E_BADLIKE)
      return efail(s, hst[w], "pofd_ens_set_positions: initial walker " + std::to_string(w) +
                                  " has no valid likelihood");
  if (pofd_ens_dev::launch_bonus_init(s->e, s->stream))
    return efail(s, POFD_E_INTERNAL, "pofd_ens_set_positions: bonus launch failed");
  (void)hipMemset(s->e.naccept, 0, W * sizeof(uint32_t));
  (void)hipMemset(s->e.nerr, 0, W * sizeof(uint32_t));
  s->nstored = 0;
  // the step counter (the random streams' index) keeps running, so new
  // positions never reuse an earlier step's draws
  s->n_exact = s->n_proposed = 0;
  if (s->e.nexact) (void)hipMemset(s->e.nexact, 0, sizeof(unsigned long long));
  s->positioned = true;
  return refresh_surrogate(s);
}

int pofd_ens_set_host_target(pofd_ens* s, pofd_host_loglike_fn fn, void* user,
                             uint32_t nbonus) {
  if (!s) return POFD_E_ARG;
  if (s->target != POFD_TARGET_HOST)
    return efail(s, POFD_E_ARG, "pofd_ens_set_host_target: the sampler's target is not POFD_TARGET_HOST");
  if (!fn) return efail(s, POFD_E_ARG, "pofd_ens_set_host_target: null function");
  if (nbonus > s->npar) return efail(s, POFD_E_ARG, "pofd_ens_set_host_target: too many bonus values");
  if (s->host_bonus_dev) {
    (void)hipFree(s->host_bonus_dev);
    s->host_bonus_dev = nullptr;
  }
  if (nbonus) {
    const hipError_t he =
        hipMalloc(&s->host_bonus_dev, static_cast<size_t>(s->nwalk) * (nbonus + 1) * sizeof(double));
    if (he != hipSuccess) return ehip(s, he, "hipMalloc(host bonus)");
  }
  s->host_fn = fn;
  s->host_user = user;
  s->host_nbonus = nbonus;
  return POFD_OK;
}

int pofd_ens_set_surrogate(pofd_ens* s, pofd_mlp* mlp, const uint32_t* param_index,
                           uint32_t nidx, double gate, double holdout_rms, int* enabled) {
  if (!s) return POFD_E_ARG;
  if (enabled) *enabled = 0;
  if (s->target == POFD_TARGET_HOST && mlp)
    return efail(s, POFD_E_ARG,
                 "pofd_ens_set_surrogate: delayed acceptance needs a device likelihood");
  s->use_da = false;
  s->mlp = mlp;
  if (!mlp) return POFD_OK;
  uint32_t nin = 0, hid = 0;
  if (pofd_mlp_info(mlp, &nin, &hid) || !param_index || nidx != nin)
    return efail(s, POFD_E_ARG, "pofd_ens_set_surrogate: input count differs from the network's");
  for (uint32_t i = 0; i < nidx; ++i)
    if (param_index[i] >= s->npar)
      return efail(s, POFD_E_ARG, "pofd_ens_set_surrogate: parameter index out of range");
    else if (s->e.state[param_index[i]] == pofd_ens_dev::kStateBonus)
      return efail(s, POFD_E_ARG, "pofd_ens_set_surrogate: bonus parameters are not inputs");
  // the held-out error gate: beyond it the plain exact step is kept
  if (!(holdout_rms <= gate)) return POFD_OK;
  const size_t W = s->nwalk, H = s->nwalk / 2, P = s->npar;
  if (!s->dabuf[0]) {
    const size_t sizes[9] = {W * sizeof(double), H * sizeof(double), H * sizeof(int32_t),
                             H * sizeof(int32_t), 64, H * P * sizeof(float),
                             H * sizeof(double), H * sizeof(int32_t), 64};
    for (int i = 0; i < 9; ++i) {
      hipError_t he = hipMalloc(&s->dabuf[i], sizes[i]);
      if (he == hipSuccess) he = hipMemset(s->dabuf[i], 0, sizes[i]);
      if (he != hipSuccess) return ehip(s, he, "hipMalloc(delayed acceptance)");
    }
    s->e.llsur = static_cast<double*>(s->dabuf[0]);
    s->e.llsp = static_cast<double*>(s->dabuf[1]);
    s->e.pass = static_cast<int32_t*>(s->dabuf[2]);
    s->e.cidx = static_cast<int32_t*>(s->dabuf[3]);
    s->e.npass = static_cast<int32_t*>(s->dabuf[4]);
    s->e.cprop = static_cast<float*>(s->dabuf[5]);
    s->e.cll = static_cast<double*>(s->dabuf[6]);
    s->e.cst = static_cast<int32_t*>(s->dabuf[7]);
    s->e.nexact = static_cast<unsigned long long*>(s->dabuf[8]);
  }

  std::vector<int> idx(param_index, param_index + nidx);
  if (s->sidx) (void)hipFree(s->sidx);
  s->sidx = nullptr;
  hipError_t he = hipMalloc(&s->sidx, nidx * sizeof(int));
  if (he == hipSuccess) he = hipMemcpy(s->sidx, idx.data(), nidx * sizeof(int), hipMemcpyHostToDevice);
  if (he != hipSuccess) return ehip(s, he, "hipMemcpy(surrogate inputs)");
  s->use_da = true;
  if (enabled) *enabled = 1;
  return refresh_surrogate(s);
}

int pofd_ens_set_surrogate_steps(pofd_ens* s, uint32_t nsteps) {
  if (!s) return POFD_E_ARG;
  if (nsteps < 1 || nsteps > kMaxInnerSteps)
    return efail(s, POFD_E_ARG, "pofd_ens_set_surrogate_steps: steps must be 1 to 64");
  if (nsteps > 1 && !s->stbuf[0]) {
    const size_t H = s->nwalk / 2, P = s->npar;
    const size_t sizes[3] = {H * P * sizeof(float), H * sizeof(double), H * sizeof(int32_t)};
    for (int i = 0; i < 3; ++i) {
      hipError_t he = hipMalloc(&s->stbuf[i], sizes[i]);
      if (he == hipSuccess) he = hipMemset(s->stbuf[i], 0, sizes[i]);
      if (he != hipSuccess) return ehip(s, he, "hipMalloc(surrogate transition)");
    }
    s->e.yin = static_cast<float*>(s->stbuf[0]);
    s->e.lsy = static_cast<double*>(s->stbuf[1]);
    s->e.moved = static_cast<int32_t*>(s->stbuf[2]);
  }
  s->st_steps = nsteps;
  return POFD_OK;
}

int pofd_ens_reset_counts(pofd_ens* s) {
  if (!s) return POFD_E_ARG;
  const size_t W = s->nwalk;
  hipError_t he = hipStreamSynchronize(s->stream);
  if (he == hipSuccess) he = hipMemset(s->e.naccept, 0, W * sizeof(uint32_t));
  if (he == hipSuccess) he = hipMemset(s->e.nerr, 0, W * sizeof(uint32_t));
  if (he != hipSuccess) return ehip(s, he, "hipMemset(counts)");
  return POFD_OK;
}

int pofd_ens_counts(pofd_ens* s, uint64_t* n_proposed, uint64_t* n_exact) {
  if (!s) return POFD_E_ARG;
  unsigned long long dev = 0;
  if (s->e.nexact) {
    hipError_t he = hipStreamSynchronize(s->stream);
    if (he == hipSuccess)
      he = hipMemcpy(&dev, s->e.nexact, sizeof(dev), hipMemcpyDeviceToHost);
    if (he != hipSuccess) return ehip(s, he, "hipMemcpy(counts)");
  }
  if (n_proposed) *n_proposed = s->n_proposed;
  if (n_exact) *n_exact = s->n_exact + dev;
  return POFD_OK;
}

int pofd_ens_run(pofd_ens* s, uint32_t nsteps, double temperature, uint32_t thin) {
  if (!s) return POFD_E_ARG;
  if (!s->positioned) return efail(s, POFD_E_ARG, "pofd_ens_run: call pofd_ens_set_positions first");
  if (!(temperature > 0.0))
    return efail(s, POFD_E_ARG, "affineEnsemble::doMasterStep: Invalid temperature");
  const uint32_t nst = (thin == 0) ? 0 : nsteps / thin;
  if (s->nstored + nst > s->capacity) {  // grow the chain storage
    const uint32_t cap = std::max(s->nstored + nst, 2 * s->capacity);
    const size_t W = s->nwalk, P = s->npar;
    float* c = nullptr;
    double* cl = nullptr;
    hipError_t he = hipMalloc(&c, static_cast<size_t>(cap) * W * P * sizeof(float));
    if (he == hipSuccess) he = hipMalloc(&cl, static_cast<size_t>(cap) * W * sizeof(double));
    if (he != hipSuccess) {
      if (c) (void)hipFree(c);
      return ehip(s, he, "hipMalloc(chain)");
    }
    if (s->nstored) {
      (void)hipStreamSynchronize(s->stream);
      (void)hipMemcpy(c, s->e.chain, static_cast<size_t>(s->nstored) * W * P * sizeof(float),
                      hipMemcpyDeviceToDevice);
      (void)hipMemcpy(cl, s->e.chainll, static_cast<size_t>(s->nstored) * W * sizeof(double),
                      hipMemcpyDeviceToDevice);
    }
    if (s->e.chain) (void)hipFree(s->e.chain);
    if (s->e.chainll) (void)hipFree(s->e.chainll);
    s->e.chain = c;
    s->e.chainll = cl;
    s->capacity = cap;
  }
  const double inv_temp = 1.0 / temperature;
  const uint32_t hw = s->nwalk / 2;
  for (uint32_t k = 0; k < nsteps; ++k) {
    const uint32_t step = s->step++;
    const int store = (thin != 0 && (k + 1) % thin == 0) ? 1 : 0;
    for (uint32_t half = 0; half < 2; ++half) {
      if (s->use_da && s->st_steps > 1) {
        s->n_proposed += hw;
        int rc = st_half(s, half, step, inv_temp, store);
        if (rc) return rc;
        continue;
      }
      if (s->use_da) {  // proposals, surrogate, both stages
        s->n_proposed += hw;
        int rc = da_half(s, half, step, inv_temp, store);
        if (rc) return rc;
        continue;
      }
      if (pofd_ens_dev::launch_propose(s->e, half, step, s->stream))
        return efail(s, POFD_E_INTERNAL, "pofd_ens_run: propose launch failed");
      s->n_proposed += hw;
      s->n_exact += hw;
      int rc = evaluate(s, s->e.prop, hw, s->e.llp, s->e.stp);
      if (rc) return efail(s, rc, "pofd_ens_run: likelihood launch failed");
      if (pofd_ens_dev::launch_accept(s->e, half, step, inv_temp, store, s->nstored,
                                      s->stream))
        return efail(s, POFD_E_INTERNAL, "pofd_ens_run: accept launch failed");
    }
    if (store) ++s->nstored;
  }
  hipError_t he = hipStreamSynchronize(s->stream);
  if (he != hipSuccess) return ehip(s, he, "hipStreamSynchronize");
  return POFD_OK;
}

int pofd_ens_get_state(pofd_ens* s, float* theta, double* ll, uint32_t* naccept,
                       uint32_t* nerr) {
  if (!s) return POFD_E_ARG;
  const size_t W = s->nwalk, P = s->npar;
  hipError_t he = hipStreamSynchronize(s->stream);
  if (he == hipSuccess && theta)
    he = hipMemcpy(theta, s->e.pos, W * P * sizeof(float), hipMemcpyDeviceToHost);
  if (he == hipSuccess && ll) he = hipMemcpy(ll, s->e.ll, W * sizeof(double), hipMemcpyDeviceToHost);
  if (he == hipSuccess && naccept)
    he = hipMemcpy(naccept, s->e.naccept, W * sizeof(uint32_t), hipMemcpyDeviceToHost);
  if (he == hipSuccess && nerr)
    he = hipMemcpy(nerr, s->e.nerr, W * sizeof(uint32_t), hipMemcpyDeviceToHost);
  if (he != hipSuccess) return ehip(s, he, "hipMemcpy(state)");
  return POFD_OK;
}

int pofd_ens_chain_length(const pofd_ens* s, uint32_t* nstored) {
  if (!s || !nstored) return POFD_E_ARG;
  *nstored = s->nstored;
  return POFD_OK;
}

int pofd_ens_get_chain(pofd_ens* s, uint32_t first, uint32_t count, float* theta, double* ll) {
  if (!s) return POFD_E_ARG;
  if (first + count > s->nstored)
    return efail(s, POFD_E_ARG, "pofd_ens_get_chain: beyond the stored steps");
  const size_t W = s->nwalk, P = s->npar;
  hipError_t he = hipStreamSynchronize(s->stream);
  if (he == hipSuccess && theta && count)
    he = hipMemcpy(theta, s->e.chain + static_cast<size_t>(first) * W * P,
                   static_cast<size_t>(count) * W * P * sizeof(float), hipMemcpyDeviceToHost);
  if (he == hipSuccess && ll && count)
    he = hipMemcpy(ll, s->e.chainll + static_cast<size_t>(first) * W,
                   static_cast<size_t>(count) * W * sizeof(double), hipMemcpyDeviceToHost);
  if (he != hipSuccess) return ehip(s, he, "hipMemcpy(chain)");
  return POFD_OK;
}

int pofd_ens_clear_chain(pofd_ens* s) {
  if (!s) return POFD_E_ARG;
  s->nstored = 0;
  return POFD_OK;
}

int pofd_ens_sample_z(pofd_ens* s, uint32_t n, double* out) {
  if (!s || !out) return POFD_E_ARG;
  double* d = nullptr;
  hipError_t he = hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(double));
  if (he != hipSuccess) return ehip(s, he, "hipMalloc(z)");
  int rc = pofd_ens_dev::launch_zsample(s->e, n, d, s->stream);
  he = hipStreamSynchronize(s->stream);
  if (he == hipSuccess && !rc) he = hipMemcpy(out, d, n * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (rc) return efail(s, POFD_E_INTERNAL, "pofd_ens_sample_z: launch failed");
  if (he != hipSuccess) return ehip(s, he, "hipMemcpy(z)");
  return POFD_OK;
}

}  // extern "C"

namespace {
// affineChainSet::acor (src/affineChainSet.cc:880-950), Goodman's ACOR:
// self-consistent window of winmult tau over maxlag = taumax * winmult
// lags, recursing on pairwise sums while tau is too long for the window.
constexpr unsigned kWinMult = 5, kTauMax = 2, kMaxLag = kTauMax * kWinMult, kMinFac = 5;

void acor_rec(double& mean, double& sigma, double& tau, std::vector<float>& X, unsigned L) {
  const double invL = 1.0 / static_cast<double>(L);
  mean = X[0];
  for (unsigned i = 1; i < L; ++i) mean += X[i];
  mean *= invL;
  for (unsigned i = 0; i < L; ++i) X[i] = static_cast<float>(X[i] - mean);
  if (L < kMinFac * kMaxLag) return;
  double C[kMaxLag + 1];
  for (unsigned s = 0; s <= kMaxLag; ++s) C[s] = 0.0;
  const unsigned iMax = L - kMaxLag;
  const double invImax = 1.0 / static_cast<double>(iMax);
  for (unsigned i = 0; i < iMax; ++i)
    for (unsigned s = 0; s <= kMaxLag; ++s) C[s] += X[i] * X[i + s];
  for (unsigned s = 0; s <= kMaxLag; ++s) C[s] *= invImax;
  double D = C[0];
  for (unsigned s = 1; s <= kMaxLag; ++s) D += 2 * C[s];
  sigma = std::sqrt(D * invL);
  tau = D / C[0];
  if (tau * kWinMult < kMaxLag) return;
  const unsigned Lh = L / 2;
  for (unsigned i = 0, j = 0; i < Lh; ++i, j += 2) X[i] = X[j] + X[j + 1];
  double newMean;
  acor_rec(newMean, sigma, tau, X, Lh);
  D = 0.25 * sigma * sigma * static_cast<double>(L);
  tau = D / C[0];
  sigma = std::sqrt(D * invL);
}
}  // namespace

extern "C" int pofd_acor(const float* x, uint32_t n, double* mean, double* sigma, double* tau) {
  if (!x || !mean || !sigma || !tau) return POFD_E_ARG;
  *sigma = *tau = NAN;
  if (n < kMinFac * kMaxLag) return POFD_E_ARG;  // affineChainSet::getAcor "succ = false"
  std::vector<float> X(x, x + n);
  acor_rec(*mean, *sigma, *tau, X, n);
  return POFD_OK;
}

extern "C" int pofd_ens_acor(pofd_ens* s, uint32_t param, double* tau, double* mean,
                             double* sigma) {
  if (!s || param >= s->npar) return POFD_E_ARG;
  const uint32_t n = s->nstored;
  const size_t W = s->nwalk, P = s->npar;
  std::vector<float> ch(static_cast<size_t>(n) * W * P);
  int rc = pofd_ens_get_chain(s, 0, n, ch.data(), nullptr);
  if (rc) return rc;
  // affineChainSet::getAverageParamVector (:830-866): walker average per step
  std::vector<float> avg(n);
  const double norm = 1.0 / static_cast<double>(W);
  for (uint32_t j = 0; j < n; ++j) {
    double cum = 0.0;
    for (size_t k = 0; k < W; ++k) cum += ch[(static_cast<size_t>(j) * W + k) * P + param];
    avg[j] = static_cast<float>(norm * cum);
  }
  return pofd_acor(avg.data(), n, mean, sigma, tau);
}
