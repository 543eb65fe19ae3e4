/* The sampler classes of the reference on the device-resident stretch move:
   affineEnsemble (include/affineEnsemble.h, src/affineEnsemble.cc) and the
   P(D) fit pofdMCMC (include/pofdMCMC.h, src/pofdMCMC.cc), driven by the
   reference's spec and init files.  Where the reference farms likelihood
   calls to MPI slaves step by step, here every half-ensemble step runs on
   the GPU (pofd_ens_*): proposals, the batched likelihood with its priors,
   acceptance, the bonus parameters and the chain.  The run logic is the
   reference's: optional initial steps at a higher temperature followed by
   a regeneration of the walkers around the best point
   (affineEnsemble::doInitialSteps, :673-740), a burn-in of min_burn steps
   extended until burn_multiple times the largest autocorrelation time
   (doBurnIn, :791-900), then nsteps stored steps per walker (masterSample,
   :600-660).  Statistics follow affineChainSet (getParamStats :1130-1200,
   getAcor :884-984).

   Header-only on top of pofd_affine_amd.hpp and pofd_affine_amd_io.hpp. */
#ifndef POFD_AFFINE_AMD_MCMC_HPP
#define POFD_AFFINE_AMD_MCMC_HPP

#include <algorithm>
#include <cmath>
#include <exception>
#include <iostream>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "pofd_affine_amd_io.hpp"

class affineEnsemble {
 public:
  /* nwalkers, nparams, nsamples (stored steps per walker after burn-in),
     init_steps, init_temp, min_burn, fixed_burn, burn_multiple, scalefac
     (affineEnsemble::affineEnsemble, src/affineEnsemble.cc:39-80) */
  affineEnsemble(unsigned int nwalkers, unsigned int nparams, unsigned int nsamples,
                 unsigned int init_steps = 0, double init_temp = 2.0,
                 unsigned int min_burn = 50, bool fixed_burn = false, float burn_multiple = 5,
                 float scalefac = 2)
      : nwalkers_(nwalkers),
        nparams_(nparams),
        nsteps_(nsamples),
        init_steps_(init_steps),
        init_temp_(init_temp),
        min_burn_(min_burn),
        fixed_burn_(fixed_burn),
        burn_multiple_(burn_multiple),
        scalefac_(scalefac) {
    if (nwalkers < 2 || (nwalkers & 1))
      throw affineExcept("affineEnsemble", "affineEnsemble",
                         "Number of walkers must be even and at least 2");
    resizeParams();
  }
  virtual ~affineEnsemble() { release(); }
  affineEnsemble(const affineEnsemble&) = delete;
  affineEnsemble& operator=(const affineEnsemble&) = delete;

  bool isValid() const noexcept { return nwalkers_ >= 2 && nparams_ > 0; }
  void setQuiet() noexcept { setVerbosity(0); }
  void setVerbose() noexcept { setVerbosity(1); }
  void setUltraVerbose() noexcept { setVerbosity(2); }
  virtual void setVerbosity(unsigned int v) noexcept { verbosity = v; }
  unsigned int getVerbosity() const noexcept { return verbosity; }
  unsigned int getNWalkers() const noexcept { return nwalkers_; }
  unsigned int getNParams() const noexcept { return nparams_; }
  void setNParams(unsigned int n) {
    nparams_ = n;
    resizeParams();
  }
  unsigned int getNSteps() const noexcept { return nsteps_; }
  float getScalefac() const noexcept { return scalefac_; }
  void setScalefac(float v) { scalefac_ = v; }
  void setSeed(unsigned long long int seed) {
    seed_ = seed;
    rangen.set_seed(seed);
  }
  /* the acceptance exponent of z: false (default) the reference's
     nparams - nfixed - 1, which counts bonus parameters
     (src/affineEnsemble.cc:1398); true the fit parameters less one, the
     exponent the stretch move's detailed balance needs */
  void setExponentMode(bool fit_dimension) { fit_dimension_ = fit_dimension; }

  /* parameter states (fixParam, setParamBonus, ignoreParamAcor,
     src/affineEnsemble.cc:230-330) */
  void clearParamState() { std::fill(state_.begin(), state_.end(), 0); }
  void fixParam(unsigned int i) { flag(i, kFixed); }
  bool isParamFixed(unsigned int i) const { return has(i, kFixed); }
  void ignoreParamAcor(unsigned int i) { flag(i, kIgnore); }
  bool isParamIgnoredAcor(unsigned int i) const { return has(i, kIgnore); }
  void setParamBonus(unsigned int i) { flag(i, kBonus); }
  bool isParamBonus(unsigned int i) const { return has(i, kBonus); }
  unsigned int getNFitParams() const noexcept { return count(kFixed | kBonus, false); }
  unsigned int getNAcorParams() const noexcept { return count(kFixed | kBonus | kIgnore, false); }
  unsigned int getNBonusParams() const noexcept { return count(kBonus, true); }
  void setParamName(unsigned int i, const std::string& n) {
    check_idx(i, "setParamName");
    names_[i] = n;
    has_name_[i] = true;
  }
  void unsetParamName(unsigned int i) {
    check_idx(i, "unsetParamName");
    has_name_[i] = false;
  }
  bool hasName(unsigned int i) const { return has_name_[i]; }
  std::string getParamName(unsigned int i) const {
    check_idx(i, "getParamName");
    return has_name_[i] ? names_[i] : std::string();
  }

  /* burn-in and sampling (affineEnsemble::sample / masterSample) */
  void sample() {
    if (!is_init) initChains();
    requireReady("sample");
    doInitialSteps();
    doBurnIn();
    runSteps(nsteps_, 1.0, true);
    if (!hasOneStepBeenAccepted())
      throw affineExcept("affineEnsemble", "masterSample",
                         "Failed to accept any steps from main loop");
  }
  /* a fixed number of stored steps (affineEnsemble::doSteps), after
     initsteps unstored ones */
  void doSteps(unsigned int n, unsigned int initsteps = 0) {
    if (!is_init) initChains();
    requireReady("doSteps");
    if (initsteps) runSteps(initsteps, 1.0, false);
    runSteps(n, 1.0, true);
  }

  /* autocorrelation times of the walker-averaged series of every parameter
     used in the burn-in check (affineChainSet::getAcor); false when a
     series is too short */
  bool computeAcor() const {
    acor_.assign(nparams_, std::numeric_limits<float>::quiet_NaN());
    bool ok = true;
    for (unsigned i = 0; i < nparams_; ++i) {
      if (state_[i] & (kFixed | kBonus | kIgnore)) continue;
      double tau = 0, mean = 0, sig = 0;
      if (pofd_ens_acor(ens_, i, &tau, &mean, &sig) != POFD_OK || !std::isfinite(tau)) {
        ok = false;
        continue;
      }
      acor_[i] = static_cast<float>(tau);
    }
    return ok;
  }
  bool getAcor(std::vector<float>& out) const {
    const bool ok = computeAcor();
    out = acor_;
    return ok;
  }
  float getMaxAcor() const {
    float m = 0.0f;
    for (float v : acor_)
      if (std::isfinite(v)) m = std::max(m, v);
    return m;
  }
  bool hasOneStepBeenAccepted() const {
    std::vector<uint32_t> na(nwalkers_);
    stateCopy(nullptr, nullptr, na.data());
    for (uint32_t v : na)
      if (v) return true;
    return false;
  }
  /* accepted steps over steps taken since the burn-in ended */
  void getAcceptanceFrac(std::vector<float>& out) const {
    std::vector<uint32_t> na(nwalkers_);
    stateCopy(nullptr, nullptr, na.data());
    out.resize(nwalkers_);
    const float n = static_cast<float>(std::max(1u, steps_since_reset_));
    for (unsigned w = 0; w < nwalkers_; ++w) out[w] = static_cast<float>(na[w]) / n;
  }
  float getParamMean(unsigned int i) const {
    float mn, var, lo, up;
    getParamStats(i, mn, var, lo, up);
    return mn;
  }
  /* mean, variance and the limits of the conflevel interval over all
     stored steps; the upper limit is read at the fraction 1 - conflevel/2
     of the sorted values, as the reference does (src/affineChainSet.cc:
     1196-1199) */
  void getParamStats(unsigned int i, float& mean, float& var, float& lowlimit, float& uplimit,
                     float conflevel = 0.683) const {
    check_idx(i, "getParamStats");
    const float nan = std::numeric_limits<float>::quiet_NaN();
    mean = var = lowlimit = uplimit = nan;
    fetchChain();
    const size_t sz = chain_ll_.size();
    if (sz == 0) return;
    std::vector<float> v(sz);
    double m = 0.0;
    for (size_t r = 0; r < sz; ++r) {
      v[r] = chain_th_[r * nparams_ + i];
      m += v[r];
    }
    const double norm = 1.0 / static_cast<double>(sz);
    m *= norm;
    mean = static_cast<float>(m);
    double sd = 0.0, s2 = 0.0;
    for (float x : v) {
      const double d = x - m;
      sd += d;
      s2 += d * d;
    }
    s2 -= sd * norm;
    s2 /= static_cast<double>(sz) - 1.0;
    var = static_cast<float>(s2);
    if (var == 0.0f) return;
    if (conflevel < 0 || conflevel > 1)
      throw affineExcept("affineChainSet", "getParamStats", "Invalid conf level");
    std::sort(v.begin(), v.end());
    const int lo = static_cast<int>(0.5 * (1.0 - conflevel) * sz);
    if (lo != 0) lowlimit = v[lo];
    const int up = static_cast<int>((1.0 - 0.5 * conflevel) * sz);
    if (up != 0) uplimit = v[up];
  }
  double getMaxLogLike() const {
    double ll;
    paramSet p(nparams_);
    getMaxLogLikeParam(ll, p);
    return ll;
  }
  void getMaxLogLikeParam(double& ll, paramSet& p) const {
    fetchChain();
    ll = -std::numeric_limits<double>::infinity();
    size_t best = 0;
    for (size_t r = 0; r < chain_ll_.size(); ++r)
      if (chain_ll_[r] > ll) {
        ll = chain_ll_[r];
        best = r;
      }
    if (chain_ll_.empty()) return;
    p = paramSet(nparams_, chain_th_.data() + best * nparams_);
  }
  virtual void printStatistics(float conflevel = 0.683, std::ostream& os = std::cout) const {
    std::ios fmt(nullptr);
    fmt.copyfmt(os);
    for (unsigned i = 0; i < nparams_; ++i) {
      const std::string name = has_name_[i] ? names_[i] : "Unknown";
      if (state_[i] & kFixed) {
        os << "Parameter: " << name << " Fixed at value: " << getParamMean(i) << std::endl;
        continue;
      }
      float mn, var, lo, up;
      getParamStats(i, mn, var, lo, up, conflevel);
      os.width(6);
      os.precision(4);
      os << "Parameter: " << name << " Mean: " << mn << " Stdev: " << std::sqrt(var) << std::endl;
      os << "  lower limit: " << lo << " upper limit: " << up << " (";
      os.width(5);
      os.precision(2);
      os << conflevel * 100.0 << "% limit)";
      if (state_[i] & kBonus) os << " (bonus param)";
      os << std::endl;
    }
    os.copyfmt(fmt);
  }
  /* the stored chain as text (affineChainSet::writeToFile) */
  void writeToFile(const std::string& path) const {
    uint32_t n = 0;
    if (ens_) pofd_ens_chain_length(ens_, &n);
    std::vector<float> th(static_cast<size_t>(n) * nwalkers_ * nparams_);
    std::vector<double> ll(static_cast<size_t>(n) * nwalkers_);
    if (n) check(pofd_ens_get_chain(ens_, 0, n, th.data(), ll.data()), "writeToFile");
    writeChainText(path, th.data(), ll.data(), n, nwalkers_, nparams_);
  }
#ifdef POFD_AFFINE_HAVE_HDF5
  /* affineEnsemble::writeToHDF5Handle / writeToHDF5,
     src/affineEnsemble.cc:1583-1767: an AffineSettings group (ScaleFactor,
     NInitSteps, MinBurn, FixedBurn, BurnMultiple, NSteps), a ParamInfo
     group (NFixed, NIgnore, NBonus, the IsParamFixed /
     IsParamAutocorrIgnore / IsParamBonus flags, InitialPosition and
     RegeneratedFirstStep when known, ParamNames when any is set) and a
     Chains group (NAccept per walker, then the chain as
     affineChainSet::writeToHDF5Handle lays it out) */
  virtual void writeToHDF5Handle(hid_t objid) const {
    using namespace pofd_hdf5;
    need(objid >= 0 && H5Iget_ref(objid) >= 0, "affineEnsemble", "writeToHDF5Handle",
         "Input handle is not valid");
    {
      Id g(createGroup(objid, "AffineSettings", "affineEnsemble", "writeToHDF5Handle"),
           H5Gclose);
      writeAttFloat(g.get(), "ScaleFactor", scalefac_);
      writeAttUnsignedInt(g.get(), "NInitSteps", init_steps_);
      writeAttUnsignedInt(g.get(), "MinBurn", min_burn_);
      writeAttBool(g.get(), "FixedBurn", fixed_burn_);
      writeAttFloat(g.get(), "BurnMultiple", burn_multiple_);
      writeAttUnsignedInt(g.get(), "NSteps", nsteps_);
    }
    {
      Id g(createGroup(objid, "ParamInfo", "affineEnsemble", "writeToHDF5Handle"), H5Gclose);
      writeAttUnsignedInt(g.get(), "NFixed", count(kFixed, true));
      writeAttUnsignedInt(g.get(), "NIgnore", count(kIgnore, true));
      writeAttUnsignedInt(g.get(), "NBonus", count(kBonus, true));
      std::vector<bool> fl(nparams_);
      for (unsigned i = 0; i < nparams_; ++i) fl[i] = (state_[i] & kFixed) != 0;
      writeAttBools(g.get(), "IsParamFixed", fl);
      for (unsigned i = 0; i < nparams_; ++i) fl[i] = (state_[i] & kIgnore) != 0;
      writeAttBools(g.get(), "IsParamAutocorrIgnore", fl);
      for (unsigned i = 0; i < nparams_; ++i) fl[i] = (state_[i] & kBonus) != 0;
      writeAttBools(g.get(), "IsParamBonus", fl);
      if (has_init_step_) writeDataFloats(g.get(), "InitialPosition", init_step_);
      if (has_regen_step_) writeDataFloats(g.get(), "RegeneratedFirstStep", regen_step_);
      if (std::find(has_name_.begin(), has_name_.end(), true) != has_name_.end())
        writeAttStrings(g.get(), "ParamNames", names_);
    }
    Id g(createGroup(objid, "Chains", "affineEnsemble", "writeToHDF5Handle"), H5Gclose);
    std::vector<uint32_t> na(nwalkers_, 0u);
    if (ens_) stateCopy(nullptr, nullptr, na.data());
    writeDataUnsignedInts(g.get(), "NAccept", std::vector<unsigned int>(na.begin(), na.end()));
    fetchChain();
    const unsigned n = static_cast<unsigned>(chain_ll_.size() / nwalkers_);
    writeChainSet(g.get(), chain_th_.data(), chain_ll_.data(), n, nwalkers_, nparams_);
  }
  void writeToHDF5(const std::string& filename) const {
    pofd_hdf5::withNewFile(filename, "affineEnsemble", "writeToHDF5",
                           [&](hid_t f) { writeToHDF5Handle(f); });
  }
#endif
  /* the stored chain, [steps][walkers][params] and [steps][walkers] */
  void getChain(std::vector<float>& theta, std::vector<double>& loglike) const {
    fetchChain();
    theta = chain_th_;
    loglike = chain_ll_;
  }

  /* the likelihood of a subclass that evaluates it on the host
     (include/affineEnsemble.h:175): used when initChains chose
     setHostTarget; pars_invalid rejects the step.  Subclasses with a
     device likelihood (pofdMCMC, pofdMCMCDouble) also answer single calls
     with it. */
  virtual double getLogLike(const paramSet&, bool& pars_invalid) {
    pars_invalid = true;
    throw affineExcept("affineEnsemble", "getLogLike",
                       "no host likelihood: override getLogLike or use a device target");
  }
  /* the bonus parameters of p after its likelihood (include/affineEnsemble.h:
     179); rej: the step was rejected.  The default leaves them alone. */
  virtual void fillBonusParams(paramSet&, bool) {}

  /* the subclass sets up its likelihood (setTarget or setHostTarget) and
     parameter states */
  virtual void initChains() = 0;
  /* the subclass draws the walkers' starting points around p
     (setInitialPositions) */
  virtual void generateInitialPosition(const paramSet& p) = 0;
  virtual bool areParamsValid(const paramSet&) const { return true; }

 protected:
  mutable ran rangen;
  bool is_init = false;
  unsigned int verbosity = 0;

  /* a standard normal deviate by the ratio-of-uniforms method of Leva
     (ACM Trans. Math. Softw. 18, 449, 1992), as ran::gauss draws it */
  double gauss() {
    double u, v, x, y, q;
    do {
      u = rangen.doub();
      v = 1.7156 * (rangen.doub() - 0.5);
      x = u - 0.449871;
      y = std::fabs(v) + 0.386595;
      q = x * x + y * (0.19600 * y - 0.25472 * x);
    } while (q > 0.27597 && (q > 0.27846 || v * v > -4.0 * std::log(u) * u * u));
    return v / u;
  }
  /* 1 + s N(0,1), redrawn while not positive (the initial sigma
     multipliers of pofdMCMC / pofdMCMCDouble) */
  double sigmaMultDraw(double s, const char* cls) {
    for (unsigned j = 0; j < 1000; ++j) {
      const double t = 1.0 + s * gauss();
      if (t > 0) return t;
    }
    throw affineExcept(cls, "generateInitialPosition", "Couldn't generate sigma multiplier value");
  }

  /* the device sampler over a likelihood context (POFD_TARGET_1D/2D), the
     parameter states and the limits (NaN: none) */
  void setTarget(pofd_ctx* ctx, int target, const std::vector<float>& lower,
                 const std::vector<float>& upper) {
    release();
    check(pofd_ens_create(&ens_, ctx, target, nwalkers_, nparams_, scalefac_, seed_ ? seed_ : 1),
          "setTarget");
    std::vector<uint8_t> st(nparams_);
    for (unsigned i = 0; i < nparams_; ++i)
      st[i] = (state_[i] & kBonus) ? 2 : ((state_[i] & kFixed) ? 1 : 0);
    check(pofd_ens_set_param_state(ens_, st.data()), "setTarget");
    check(pofd_ens_set_limits(ens_, lower.empty() ? nullptr : lower.data(),
                              upper.empty() ? nullptr : upper.data()),
          "setTarget");
    float dexp = 0.0f;
    check(pofd_ens_set_exponent_mode(ens_, fit_dimension_ ? 1 : 0, &dexp), "setTarget");
  }
  /* the sampler over the subclass's own getLogLike / fillBonusParams on the
     host: proposals and acceptance stay on the device, each half-step's
     proposals come to the host for their likelihoods (a proposal for which
     areParamsValid is false is rejected without a likelihood call) */
  void setHostTarget(const std::vector<float>& lower = std::vector<float>(),
                     const std::vector<float>& upper = std::vector<float>()) {
    setTarget(nullptr, POFD_TARGET_HOST, lower, upper);
    check(pofd_ens_set_host_target(ens_, &affineEnsemble::hostLoglike, this, getNBonusParams()),
          "setHostTarget");
  }
  /* the starting point the walkers were drawn around (recorded as
     ParamInfo/InitialPosition, src/pofdMCMC.cc:205-206) */
  void setInitStep(const paramSet& p) {
    has_init_step_ = true;
    init_step_.assign(p.data(), p.data() + p.getNParams());
  }
  /* one position per walker; evaluates them (and their bonus parameters) */
  void setInitialPositions(const std::vector<paramSet>& pos) {
    if (pos.size() != nwalkers_)
      throw affineExcept("affineEnsemble", "generateInitialPosition", "One position per walker");
    std::vector<float> th(static_cast<size_t>(nwalkers_) * nparams_);
    for (unsigned w = 0; w < nwalkers_; ++w)
      for (unsigned i = 0; i < nparams_; ++i) th[static_cast<size_t>(w) * nparams_ + i] = pos[w][i];
    check(pofd_ens_set_positions(ens_, th.data()), "generateInitialPosition");
    steps_since_reset_ = 0;
    chain_valid_ = false;
  }

 private:
  enum { kFixed = 1, kIgnore = 2, kBonus = 4 };
  static constexpr unsigned kMinAcorExtra = 10, kMaxAcorIters = 20, kMaxBurnIters = 30;
  unsigned int nwalkers_, nparams_, nsteps_, init_steps_;
  double init_temp_;
  unsigned int min_burn_;
  bool fixed_burn_;
  float burn_multiple_, scalefac_;
  unsigned long long seed_ = 0;
  bool fit_dimension_ = false;
  std::vector<int> state_;
  std::vector<std::string> names_;
  std::vector<bool> has_name_;
  pofd_ens* ens_ = nullptr;
  mutable std::vector<float> acor_;
  mutable std::vector<float> chain_th_;
  mutable std::vector<double> chain_ll_;
  mutable bool chain_valid_ = false;
  unsigned int steps_since_reset_ = 0;
  bool has_init_step_ = false, has_regen_step_ = false;
  std::vector<float> init_step_, regen_step_;

  void resizeParams() {
    state_.assign(nparams_, 0);
    names_.assign(nparams_, std::string());
    has_name_.assign(nparams_, false);
  }
  void check_idx(unsigned i, const char* m) const {
    if (i >= nparams_) throw affineExcept("affineEnsemble", m, "Invalid parameter index");
  }
  void flag(unsigned i, int f) {
    check_idx(i, "setParamState");
    state_[i] |= f;
  }
  bool has(unsigned i, int f) const {
    check_idx(i, "getParamState");
    return (state_[i] & f) != 0;
  }
  unsigned count(int mask, bool any) const {
    unsigned n = 0;
    for (int s : state_) n += any ? ((s & mask) != 0) : ((s & mask) == 0);
    return n;
  }
  std::exception_ptr host_error_;
  static int hostLoglike(void* user, const float* th, uint32_t n, uint32_t stride, double* ll,
                         int32_t* st, double* bonus) {
    affineEnsemble* self = static_cast<affineEnsemble*>(user);
    try {
      const unsigned P = self->nparams_;
      std::vector<unsigned> bidx;
      for (unsigned i = 0; i < P; ++i)
        if (self->state_[i] & kBonus) bidx.push_back(i);
      for (uint32_t r = 0; r < n; ++r) {
        paramSet p(P, th + static_cast<size_t>(r) * stride);
        bool rej = false;
        double v = 0.0;
        if (!self->areParamsValid(p))
          rej = true;
        else
          v = self->getLogLike(p, rej);
        ll[r] = v;
        st[r] = rej ? POFD_E_ARG : POFD_OK;
        if (bonus && !bidx.empty()) {
          self->fillBonusParams(p, rej);
          for (size_t b = 0; b < bidx.size(); ++b) bonus[r * bidx.size() + b] = p[bidx[b]];
        }
      }
      return POFD_OK;
    } catch (...) {
      self->host_error_ = std::current_exception();
      return POFD_E_INTERNAL;
    }
  }
  void check(int rc, const char* m) const {
    if (rc && host_error_) {
      std::exception_ptr e = host_error_;
      const_cast<affineEnsemble*>(this)->host_error_ = nullptr;
      std::rethrow_exception(e);
    }
    if (rc) {
      const char* e = ens_ ? pofd_ens_last_error(ens_) : "";
      throw affineExcept("affineEnsemble", m, std::string(e && *e ? e : "device sampler error"));
    }
  }
  void release() {
    if (ens_) pofd_ens_destroy(ens_);
    ens_ = nullptr;
  }
  void requireReady(const char* m) const {
    if (!ens_) throw affineExcept("affineEnsemble", m, "initChains did not create the sampler");
  }
  void stateCopy(float* th, double* ll, uint32_t* na) const {
    if (!ens_) throw affineExcept("affineEnsemble", "getState", "No sampler");
    check(pofd_ens_get_state(ens_, th, ll, na, nullptr), "getState");
  }
  void fetchChain() const {
    if (chain_valid_) return;
    uint32_t n = 0;
    if (ens_) pofd_ens_chain_length(ens_, &n);
    chain_th_.assign(static_cast<size_t>(n) * nwalkers_ * nparams_, 0.0f);
    chain_ll_.assign(static_cast<size_t>(n) * nwalkers_, 0.0);
    if (n) check(pofd_ens_get_chain(ens_, 0, n, chain_th_.data(), chain_ll_.data()), "getChain");
    chain_valid_ = true;
  }
  void runSteps(unsigned n, double temp, bool store) {
    check(pofd_ens_run(ens_, n, temp, store ? 1u : 0u), "doMasterStep");
    steps_since_reset_ += n;
    chain_valid_ = false;
  }
  /* init_steps at init_temp, then the walkers regenerated around the best
     point seen (doInitialSteps) */
  void doInitialSteps() {
    if (init_steps_ == 0) return;
    check(pofd_ens_clear_chain(ens_), "doInitialSteps");
    runSteps(init_steps_, init_temp_, true);
    if (!hasOneStepBeenAccepted())
      throw affineExcept("affineEnsemble", "doInitialSteps", "Failed to accept any initial steps");
    double ll;
    paramSet best(nparams_);
    getMaxLogLikeParam(ll, best);
    generateInitialPosition(best);
    has_regen_step_ = true;
    regen_step_.assign(best.data(), best.data() + best.getNParams());
  }
  /* min_burn steps, extended until the stored length exceeds
     burn_multiple x the largest autocorrelation time (doBurnIn,
     addStepsToGetAcor); the burn-in is then discarded */
  void doBurnIn() {
    check(pofd_ens_clear_chain(ens_), "doBurnIn");
    runSteps(min_burn_, 1.0, true);
    if (!hasOneStepBeenAccepted())
      throw affineExcept("affineEnsemble", "doBurnIn",
                         "Failed to accept any steps from initial burn in");
    unsigned nsteps = min_burn_;
    if (!fixed_burn_) {
      nsteps += addStepsToGetAcor();
      unsigned nmin = static_cast<unsigned>(burn_multiple_ * getMaxAcor() + 0.999999999999999);
      for (unsigned it = 0; nsteps <= nmin; ++it) {
        if (it > kMaxBurnIters)
          throw affineExcept("affineEnsemble", "doBurnIn", "Failed to converge in burn in");
        const unsigned more = (nmin - nsteps) / 2 + 1;
        runSteps(more, 1.0, true);
        nsteps += more;
        nsteps += addStepsToGetAcor();
        nmin = static_cast<unsigned>(burn_multiple_ * getMaxAcor() + 0.999999999999999);
      }
    }
    check(pofd_ens_clear_chain(ens_), "doBurnIn");
    check(pofd_ens_reset_counts(ens_), "doBurnIn");  // acceptance restarts
    steps_since_reset_ = 0;
    chain_valid_ = false;
  }
  unsigned addStepsToGetAcor() {
    if (computeAcor()) return 0;
    const unsigned extra = std::max(static_cast<unsigned>(min_burn_ * 0.25), kMinAcorExtra);
    unsigned taken = 0;
    for (unsigned i = 0; i < kMaxAcorIters; ++i) {
      runSteps(extra, 1.0, true);
      taken += extra;
      if (computeAcor()) return taken;
    }
    throw affineExcept("affineEnsemble", "addToGetAcor", "Can't compute acor; increase min_burn");
  }
};

/* ---- the 1D P(D) fit (pofdMCMC, src/pofdMCMC.cc): knots from the init
   file (ranges, limits, fixed knots), the likelihood set up from the spec
   file, the sigma multiplier fitted when the spec asks for it, and two
   bonus parameters, <S>area and <S^2>area. ---- */
class pofdMCMC : public affineEnsemble {
 public:
  pofdMCMC(const std::string& initfile, const std::string& specfile, unsigned int nwalkers,
           unsigned int nsamples, unsigned int init_steps = 0, double init_temp = 2.0,
           unsigned int min_burn = 50, bool fixed_burn = false, float burn_multiple = 5,
           float scalefac = 2)
      : affineEnsemble(nwalkers, 1, nsamples, init_steps, init_temp, min_burn, fixed_burn,
                       burn_multiple, scalefac),
        initfile_(initfile),
        specfile_(specfile) {}

  /* pofdMCMC::initChainsMaster (src/pofdMCMC.cc:115-215) */
  void initChains() override {
    spec_info.readFile(specfile_);
    if (spec_info.datafiles.empty())
      throw affineExcept("pofdMCMC", "initChainsMaster", "No data files read");
    ifile.readFile(initfile_, true, true);
    nknots_ = ifile.getNKnots();
    const unsigned npar = nknots_ + 3;
    setNParams(npar);
    for (unsigned i = 0; i < nknots_; ++i) {
      setParamName(i, "Knot" + std::to_string(i));
      if (ifile.isKnotFixed(i)) fixParam(i);
    }
    setParamName(nknots_, "SigmaMult");
    setParamName(nknots_ + 1, "<S>area");
    setParamName(nknots_ + 2, "<S^2>area");
    if (!spec_info.fit_sigma) fixParam(nknots_);
    setParamBonus(nknots_ + 1);
    setParamBonus(nknots_ + 2);
    likeSet = std::make_unique<calcLike>(spec_info.fftsize, spec_info.ninterp, spec_info.bin_data,
                                         spec_info.nbins);
    if (spec_info.has_cfirbprior)
      likeSet->setCFIRBPrior(spec_info.cfirbprior_mean, spec_info.cfirbprior_stdev);
    if (spec_info.has_poissonprior)
      likeSet->setPoissonPrior(spec_info.poissonprior_mean, spec_info.poissonprior_stdev);
    if (spec_info.fit_sigma && spec_info.has_sigprior)
      likeSet->setSigmaPrior(spec_info.sigprior_stdev);
    if (spec_info.regularization_alpha > 0.0)
      likeSet->setRegularizationAlpha(spec_info.regularization_alpha);
    likeSet->readDataFromFiles(spec_info.datafiles, spec_info.psffiles, spec_info.sigmas,
                               spec_info.like_norm, spec_info.ignore_mask, spec_info.mean_sub,
                               spec_info.minbeamval, spec_info.beam_histogram,
                               spec_info.nbeamhist, spec_info.exp_conf);
    std::vector<double> kp;
    ifile.getKnotPos(kp);
    likeSet->setKnotPositions(kp);
    paramSet p(npar);
    ifile.getParams(p);
    p[nknots_] = 1.0f;
    p[nknots_ + 1] = p[nknots_ + 2] = std::numeric_limits<float>::quiet_NaN();
    likeSet->setRRanges(p);
    setVerbosity(spec_info.verbosity);
    if (spec_info.has_user_seed) setSeed(spec_info.seed);
    // limits of the knots from the init file; the sigma multiplier > 0
    const float nan = std::numeric_limits<float>::quiet_NaN();
    std::vector<float> lo(npar, nan), up(npar, nan);
    for (unsigned i = 0; i < nknots_; ++i) {
      if (ifile.knotHasLowerLimit(i)) lo[i] = static_cast<float>(ifile.getLowerLimit(i));
      if (ifile.knotHasUpperLimit(i)) up[i] = static_cast<float>(ifile.getUpperLimit(i));
    }
    lo[nknots_] = std::numeric_limits<float>::min();
    setTarget(likeSet->context(), POFD_TARGET_1D, lo, up);
    is_init = true;
    setInitStep(p);
    generateInitialPosition(p);
  }

  /* pofdMCMC::areParamsValid (src/pofdMCMC.cc:216-224) */
  bool areParamsValid(const paramSet& p) const override {
    if (!ifile.isValid(p)) return false;
    return p[nknots_] > 0.0f;
  }

  /* pofdMCMC::generateInitialPosition (src/pofdMCMC.cc:226-292): knots
     drawn in their init-file ranges around p, the sigma multiplier as
     1 + s N(0,1) with s the sigma prior width (or 0.1), redrawn while not
     positive */
  void generateInitialPosition(const paramSet& p) override {
    const unsigned npar = getNParams();
    if (p.getNParams() < npar)
      throw affineExcept("pofdMCMC", "generateInitialPosition", "Wrong number of params in input");
    const double sm = spec_info.has_sigprior ? spec_info.sigprior_stdev : 0.1;
    std::vector<paramSet> pos;
    for (unsigned w = 0; w < getNWalkers(); ++w) {
      paramSet q(npar);
      ifile.generateRandomKnotValues(rangen, q, p);
      q[nknots_] = spec_info.fit_sigma ? static_cast<float>(sigmaMultDraw(sm, "pofdMCMC")) : 1.0f;
      q[nknots_ + 1] = q[nknots_ + 2] = std::numeric_limits<float>::quiet_NaN();
      pos.push_back(q);
    }
    setInitialPositions(pos);
  }

  /* a single log likelihood (pofdMCMC::getLogLike), priors included */
  double getLogLike(const paramSet& p, bool& rejected) override {
    return likeSet->getLogLike(p, rejected);
  }

#ifdef POFD_AFFINE_HAVE_HDF5
  /* pofdMCMC::writeToHDF5Handle, src/pofdMCMC.cc:365-393: the ensemble's
     groups, the init file's ranges and limits added to ParamInfo, and a
     LikelihoodParams group with the likelihood's and the spec file's
     settings */
  void writeToHDF5Handle(hid_t objid) const override {
    using namespace pofd_hdf5;
    affineEnsemble::writeToHDF5Handle(objid);
    {
      Id g(H5Gopen2(objid, "ParamInfo", H5P_DEFAULT), H5Gclose);
      need(g.ok(), "pofdMCMC", "writeToHDF5Handle", "Can't open ParamInfo group");
      ifile.writeToHDF5Handle(g.get());
    }
    Id g(createGroup(objid, "LikelihoodParams", "pofdMCMC", "writeToHDF5Handle"), H5Gclose);
    if (likeSet) likeSet->writeToHDF5Handle(g.get());
    spec_info.writeToHDF5Handle(g.get());
  }
#endif

  specFile spec_info;
  initFileKnots ifile;

 private:
  std::string initfile_, specfile_;
  unsigned nknots_ = 0;
  std::unique_ptr<calcLike> likeSet;

};

/* ---- the 2D P(D) fit (pofdMCMCDouble, src/pofdMCMCDouble.cc): band-1
   knots, colour sigma and offset knots from the init file, the likelihood
   from the spec file, two sigma multipliers, and four bonus parameters
   (<S1>area, <S2>area, <S1^2>area, <S2^2>area). ---- */
class pofdMCMCDouble : public affineEnsemble {
 public:
  pofdMCMCDouble(const std::string& initfile, const std::string& specfile, unsigned int nwalkers,
                 unsigned int nsamples, unsigned int init_steps = 0, double init_temp = 2.0,
                 unsigned int min_burn = 50, bool fixed_burn = false, float burn_multiple = 5,
                 float scalefac = 2)
      : affineEnsemble(nwalkers, 1, nsamples, init_steps, init_temp, min_burn, fixed_burn,
                       burn_multiple, scalefac),
        initfile_(initfile),
        specfile_(specfile) {}

  /* pofdMCMCDouble::initChainsMaster (src/pofdMCMCDouble.cc) */
  void initChains() override {
    spec_info.readFile(specfile_);
    if (spec_info.datafiles1.empty())
      throw affineExcept("pofdMCMCDouble", "initChainsMaster", "No data files read");
    ifile.readFile(initfile_, true, true);
    ntot_ = ifile.getNTot();
    const unsigned npar = ntot_ + 6;
    setNParams(npar);
    const unsigned nk = ifile.getNKnots(), ns = ifile.getNSigmas();
    for (unsigned i = 0; i < ntot_; ++i) {
      if (i < nk) setParamName(i, "Knot" + std::to_string(i));
      else if (i < nk + ns) setParamName(i, "Sigma" + std::to_string(i - nk));
      else setParamName(i, "Offset" + std::to_string(i - nk - ns));
      if (ifile.isKnotFixed(i)) fixParam(i);
    }
    const char* names[6] = {"SigmaMult1", "SigmaMult2", "<S1>area",
                            "<S2>area",   "<S1^2>area", "<S2^2>area"};
    for (unsigned i = 0; i < 6; ++i) setParamName(ntot_ + i, names[i]);
    if (!spec_info.fit_sigma1) fixParam(ntot_);
    if (!spec_info.fit_sigma2) fixParam(ntot_ + 1);
    for (unsigned i = ntot_ + 2; i < npar; ++i) setParamBonus(i);
    likeSet = std::make_unique<calcLikeDouble>(spec_info.fftsize, spec_info.nedge,
                                               spec_info.edge_set, spec_info.bin_data,
                                               spec_info.nbins);
    if (spec_info.has_cfirbprior1)
      likeSet->setCFIRBPrior1(spec_info.cfirbprior_mean1, spec_info.cfirbprior_stdev1);
    if (spec_info.has_cfirbprior2)
      likeSet->setCFIRBPrior2(spec_info.cfirbprior_mean2, spec_info.cfirbprior_stdev2);
    if (spec_info.has_poissonprior1)
      likeSet->setPoissonPrior1(spec_info.poissonprior_mean1, spec_info.poissonprior_stdev1);
    if (spec_info.has_poissonprior2)
      likeSet->setPoissonPrior2(spec_info.poissonprior_mean2, spec_info.poissonprior_stdev2);
    if (spec_info.fit_sigma1 && spec_info.has_sigprior1)
      likeSet->setSigmaPrior1(spec_info.sigprior_stdev1);
    if (spec_info.fit_sigma2 && spec_info.has_sigprior2)
      likeSet->setSigmaPrior2(spec_info.sigprior_stdev2);
    if (spec_info.regularization_alpha > 0.0)
      likeSet->setRegularizationAlpha(spec_info.regularization_alpha);
    likeSet->readDataFromFiles(spec_info.datafiles1, spec_info.datafiles2, spec_info.psffiles1,
                               spec_info.psffiles2, spec_info.sigmas1, spec_info.sigmas2,
                               spec_info.like_norm, spec_info.ignore_mask, spec_info.mean_sub,
                               spec_info.minbeamval, spec_info.beam_histogram,
                               spec_info.nbeamhist, spec_info.exp_conf1, spec_info.exp_conf2);
    std::vector<double> K, S, O;
    ifile.getKnotPos(K);
    ifile.getSigmaPos(S);
    ifile.getOffsetPos(O);
    likeSet->setPositions(K, S, O);
    paramSet p(npar);
    ifile.getParams(p);
    p[ntot_] = p[ntot_ + 1] = 1.0f;
    for (unsigned i = ntot_ + 2; i < npar; ++i) p[i] = std::numeric_limits<float>::quiet_NaN();
    likeSet->setRRanges(p);
    setVerbosity(spec_info.verbosity);
    if (spec_info.has_user_seed) setSeed(spec_info.seed);
    const float nan = std::numeric_limits<float>::quiet_NaN();
    std::vector<float> lo(npar, nan), up(npar, nan);
    for (unsigned i = 0; i < ntot_; ++i) {
      if (ifile.knotHasLowerLimit(i)) lo[i] = static_cast<float>(ifile.getLowerLimit(i));
      if (ifile.knotHasUpperLimit(i)) up[i] = static_cast<float>(ifile.getUpperLimit(i));
    }
    lo[ntot_] = lo[ntot_ + 1] = std::numeric_limits<float>::min();
    setTarget(likeSet->context(), POFD_TARGET_2D, lo, up);
    is_init = true;
    setInitStep(p);
    generateInitialPosition(p);
  }

  bool areParamsValid(const paramSet& p) const override {
    if (!ifile.isValid(p)) return false;
    return p[ntot_] > 0.0f && p[ntot_ + 1] > 0.0f;
  }

  /* pofdMCMCDouble::generateInitialPosition: knots, sigmas and offsets in
     their init-file ranges around p, each sigma multiplier 1 + s N(0,1)
     (s its prior width, or 0.1) when fitted */
  void generateInitialPosition(const paramSet& p) override {
    const unsigned npar = getNParams();
    if (p.getNParams() < npar)
      throw affineExcept("pofdMCMCDouble", "generateInitialPosition",
                         "Wrong number of params in input");
    const double s1 = spec_info.has_sigprior1 ? spec_info.sigprior_stdev1 : 0.1;
    const double s2 = spec_info.has_sigprior2 ? spec_info.sigprior_stdev2 : 0.1;
    std::vector<paramSet> pos;
    for (unsigned w = 0; w < getNWalkers(); ++w) {
      paramSet q(npar);
      ifile.generateRandomKnotValues(rangen, q, p);
      q[ntot_] = spec_info.fit_sigma1 ? static_cast<float>(sigmaMultDraw(s1, "pofdMCMCDouble"))
                                      : 1.0f;
      q[ntot_ + 1] =
          spec_info.fit_sigma2 ? static_cast<float>(sigmaMultDraw(s2, "pofdMCMCDouble")) : 1.0f;
      for (unsigned i = ntot_ + 2; i < npar; ++i) q[i] = std::numeric_limits<float>::quiet_NaN();
      pos.push_back(q);
    }
    setInitialPositions(pos);
  }

  double getLogLike(const paramSet& p, bool& rejected) override {
    return likeSet->getLogLike(p, rejected);
  }

  specFileDouble spec_info;
  initFileDoubleLogNormal ifile;

 private:
  std::string initfile_, specfile_;
  unsigned ntot_ = 0;
  std::unique_ptr<calcLikeDouble> likeSet;
};

#endif  // POFD_AFFINE_AMD_MCMC_HPP
