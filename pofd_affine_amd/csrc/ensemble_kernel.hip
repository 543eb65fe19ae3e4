// Device-resident affine-invariant ensemble sampler (the stretch move of
// affineEnsemble, src/affineEnsemble.cc:116-126, 1144-1420): proposals,
// acceptance and chain storage run on the GPU between the batched
// likelihood launches, so a whole run needs no host round trip per step.
//
//   propose   for every walker of one half: a partner drawn uniformly from
//             the other half (ran::selectFromRange, src/ran.cc:37-41),
//             z ~ g(z) ~ 1/sqrt(z) on [1/a, a] (generateZ, :116-126),
//             new = z old + (1 - z) partner on the fit parameters, fixed
//             parameters kept, bonus parameters NaN, and up to 40 redraws of
//             z while the proposal leaves the parameter limits
//             (generateNewStep, :1144-1228)
//   accept    prob = exp((nparams - nfixed - 1) ln z + (L_new - L_old)/T),
//             accept when prob >= 1 or u < prob (emptyMasterQueue,
//             :1398-1420); the walker's current state and the stored chain
//             are updated
//
// Random numbers: Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11), keyed by the run seed and
// counted by (step, half, walker, draw), so every run is reproducible and
// independent of the launch shape.
#include <hip/hip_runtime.h>

#include "ensemble_internal.h"
#include "mlp_infer.h"

namespace pofd_ens_dev {

namespace {

__device__ __forceinline__ void philox_round(uint32_t& c0, uint32_t& c1, uint32_t& c2,
                                             uint32_t& c3, uint32_t k0, uint32_t k1) {
  const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
  const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
  const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
  const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
  c0 = hi1 ^ c1 ^ k0;
  c1 = lo1;
  c2 = hi0 ^ c3 ^ k1;
  c3 = lo0;
}

// four 32-bit words for counter (a, b, c, d) under key (k0, k1)
__device__ __forceinline__ uint4 philox(uint32_t a, uint32_t b, uint32_t c, uint32_t d,
                                        uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    philox_round(a, b, c, d, k0, k1);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return make_uint4(a, b, c, d);
}

// uniform double in [0, 1) from 53 random bits
__device__ __forceinline__ double u53(uint32_t hi, uint32_t lo) {
  const uint64_t v = (static_cast<uint64_t>(hi) << 21) ^ (static_cast<uint64_t>(lo) >> 11);
  return static_cast<double>(v) * 1.1102230246251565404e-16;  // 2^-53
}

struct Draws {
  double u0, u1;
};

__device__ __forceinline__ Draws draws(const EnsDev& e, uint32_t step, uint32_t half,
                                       uint32_t walker, uint32_t draw) {
  const uint4 r = philox(step, (half << 16) ^ draw, walker, 0x50F0D00Du,
                         static_cast<uint32_t>(e.seed), static_cast<uint32_t>(e.seed >> 32));
  return Draws{u53(r.x, r.y), u53(r.z, r.w)};
}

// affineEnsemble::generateZ: G^{-1}(x) = ((a-1) x + 1)^2 / a
__device__ __forceinline__ float gen_z(double a, double u) {
  const double val = (a - 1.0) * u + 1.0;
  return static_cast<float>(val * val / a);
}

__device__ __forceinline__ bool params_valid(const EnsDev& e, const float* p) {
  for (int i = 0; i < e.npar; ++i) {
    if (e.state[i] != kStateFit) continue;
    const float v = p[i];
    if (!(v == v)) return false;  // NaN
    if (e.has_lower[i] && v < e.lower[i]) return false;
    if (e.has_upper[i] && v > e.upper[i]) return false;
  }
  return true;
}

// the bonus parameters of walker position x from batch row r
__device__ __forceinline__ void fill_bonus(const EnsDev& e, float* x, int r) {
  int nb = 0;
  for (int i = 0; i < e.npar; ++i) {
    if (e.state[i] != kStateBonus) continue;
    x[i] = (e.bonus_src && nb < e.bonus_count)
               ? static_cast<float>(
                     e.bonus_src[static_cast<size_t>(r) * e.bonus_stride + 1 + nb])
               : __int_as_float(0x7fc00000);
    ++nb;
  }
}

}  // namespace

// the proposal of walker t of half `half`: its position is its current
// one, or row t of `from` (the surrogate-transition inner chain); its
// draws are numbered from draw0
__device__ void propose_one(const EnsDev& e, const uint32_t half, const uint32_t step,
                            const int t, const float* from, const uint32_t draw0) {
  const int hw = e.nwalk / 2;
  const int k = static_cast<int>(half) * hw + t;       // walker updated
  const int other = (half == 0) ? hw : 0;              // partner half
  const Draws d0 = draws(e, step, half, k, draw0);
  int j = other + static_cast<int>(static_cast<double>(hw) * d0.u0);
  if (j >= other + hw) j = other + hw - 1;
  const float* xo = from ? from + static_cast<size_t>(t) * e.npar
                         : e.pos + static_cast<size_t>(k) * e.npar;
  const float* xp = e.pos + static_cast<size_t>(j) * e.npar;
  float* y = e.prop + static_cast<size_t>(t) * e.npar;
  float z = gen_z(e.scalefac, d0.u1);
  bool ok = false;
  for (int it = 0; it <= 41 && !ok; ++it) {  // first try + maxiters redraws
    if (it > 0)
      z = gen_z(e.scalefac, draws(e, step, half, k, draw0 + static_cast<uint32_t>(it)).u0);
    const float omz = 1.0f - z;
    for (int i = 0; i < e.npar; ++i) {
      const int st = e.state[i];
      if (st == kStateFixed)
        y[i] = xo[i];
      else if (st == kStateBonus)
        y[i] = __int_as_float(0x7fc00000);
      else
        y[i] = z * xo[i] + omz * xp[i];
    }
    ok = params_valid(e, y);
  }
  e.z[t] = z;
  e.pvalid[t] = ok ? 1 : 0;
}

// one thread per walker of half `half`
__global__ __launch_bounds__(64) void propose_kernel(const EnsDev e, const uint32_t half,
                                                     const uint32_t step, const float* from,
                                                     const uint32_t draw0) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < e.nwalk / 2) propose_one(e, half, step, t, from, draw0);
}

// diagonal Gaussian test target: L = -1/2 sum ((x - mu)/sigma)^2 over the
// fit parameters
__global__ __launch_bounds__(64) void gauss_target_kernel(const EnsDev e, const float* x,
                                                          const int n, double* ll,
                                                          int32_t* st) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float* p = x + static_cast<size_t>(t) * e.npar;
  double s = 0.0;
  for (int i = 0; i < e.npar; ++i) {
    if (e.state[i] != kStateFit) continue;
    const double r = (static_cast<double>(p[i]) - e.gmean[i]) * e.gisig[i];
    s += r * r;
  }
  ll[t] = -0.5 * s;
  st[t] = 0;
}

__global__ __launch_bounds__(64) void accept_kernel(const EnsDev e, const uint32_t half,
                                                    const uint32_t step, const double inv_temp,
                                                    const int store, const uint32_t slot) {
  const int hw = e.nwalk / 2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= hw) return;
  const int k = static_cast<int>(half) * hw + t;
  float* xo = e.pos + static_cast<size_t>(k) * e.npar;
  const float* y = e.prop + static_cast<size_t>(t) * e.npar;
  bool accept = false;
  if (!e.pvalid[t]) {
    e.nerr[k] += 1;  // no valid proposal after the redraws
  } else if (e.stp[t] != 0 && e.stp[t] != kBadLike) {
    e.nerr[k] += 1;  // the likelihood failed (split, model, mean)
  } else {
    const double lnew = e.llp[t], lold = e.ll[k];
    if (isnan(lnew) || isinf(lnew)) {
      e.nerr[k] += 1;
    } else {
      const double prob = exp(static_cast<double>(e.dexp) * log(static_cast<double>(e.z[t])) +
                              inv_temp * (lnew - lold));
      if (prob >= 1.0) {
        accept = true;
      } else {
        const double u = draws(e, step, half, k, 100).u0;
        accept = u < prob;
      }
    }
  }
  if (accept) {
    for (int i = 0; i < e.npar; ++i) xo[i] = y[i];
    fill_bonus(e, xo, t);
    e.ll[k] = e.llp[t];
    e.naccept[k] += 1;
  }
  if (store) {
    float* c = e.chain + (static_cast<size_t>(slot) * e.nwalk + k) * e.npar;
    for (int i = 0; i < e.npar; ++i) c[i] = xo[i];
    e.chainll[static_cast<size_t>(slot) * e.nwalk + k] = e.ll[k];
  }
}

// ---- delayed acceptance (Christen and Fox, J. Comput. Graph. Stat. 14,
// 795 (2005)): the stretch move's acceptance split into a first stage on the
// surrogate L~ and a second, exact stage that corrects it, so the chain
// keeps the exact posterior while proposals the surrogate rejects never
// reach the exact likelihood.
//   stage 1: a1 = z^d exp((L~new - L~old)/T)
//   stage 2: a2 = exp(((Lnew - Lold) - (L~new - L~old))/T)
// First stage and compaction in one workgroup: pass/fail per proposal,
// then an order-preserving scan (wave ballots, wave offsets in LDS) that
// lists the survivors in cidx, counts them in npass, records each
// proposal's compact index in pass (-1: rejected) and copies the
// survivors' rows of ys into cprop (each by the thread that placed it).
constexpr int kCompactThreads = 1024;

// moved_only: no first stage; the walkers whose surrogate-transition inner
// chain moved are the survivors.  Run by all kCompactThreads threads of
// one workgroup.
__device__ void stage1_body(const EnsDev& e, const uint32_t half, const uint32_t step,
                            const double inv_temp, const int moved_only, const float* ys) {
  __shared__ int woff[kCompactThreads / 64 + 1];
  const int hw = e.nwalk / 2;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int base = 0;
  for (int t0 = 0; t0 < hw; t0 += kCompactThreads) {
    const int t = t0 + tid;
    int pass = 0;
    if (t < hw && moved_only) {
      pass = e.moved[t];
    } else if (t < hw) {
      const int k = static_cast<int>(half) * hw + t;
      if (!e.pvalid[t]) {
        e.nerr[k] += 1;
      } else {
        const double ls = e.llsp[t];
        if (!(isnan(ls) || isinf(ls))) {
          const double a1 = exp(static_cast<double>(e.dexp) * log(static_cast<double>(e.z[t])) +
                                inv_temp * (ls - e.llsur[k]));
          pass = (a1 >= 1.0) || (draws(e, step, half, k, 101).u0 < a1);
        }
      }
    }
    const unsigned long long bal = __ballot(pass);
    if (lane == 0) woff[wave] = __popcll(bal);
    __syncthreads();
    if (tid == 0) {
      int acc = 0;
      for (int w = 0; w < kCompactThreads / 64; ++w) {
        const int c = woff[w];
        woff[w] = acc;
        acc += c;
      }
      woff[kCompactThreads / 64] = acc;
    }
    __syncthreads();
    if (t < hw) {
      int c = -1;
      if (pass) {
        const unsigned long long below = (lane == 0) ? 0ull : (bal & (~0ull >> (64 - lane)));
        c = base + woff[wave] + __popcll(below);
        e.cidx[c] = t;
        const float* src = ys + static_cast<size_t>(t) * e.npar;
        float* dst = e.cprop + static_cast<size_t>(c) * e.npar;
        for (int i = 0; i < e.npar; ++i) dst[i] = src[i];
      }
      e.pass[t] = c;
    }
    base += woff[kCompactThreads / 64];
    __syncthreads();
  }
  if (tid == 0) {
    *e.npass = base;
    *e.nexact += static_cast<unsigned long long>(base);
  }
}

__global__ __launch_bounds__(kCompactThreads) void stage1_kernel(const EnsDev e,
                                                                 const uint32_t half,
                                                                 const uint32_t step,
                                                                 const double inv_temp,
                                                                 const int moved_only,
                                                                 const float* ys) {
  stage1_body(e, half, step, inv_temp, moved_only, ys);
}

// one thread per walker t of the half: a survivor (compact index pass[t])
// is accepted or not on its exact log likelihood (row pass[t] of cll),
// its candidate row t of ys with surrogate lsn[t]; then, when storing, the
// walker's state goes into chain slot `slot`
__global__ __launch_bounds__(64) void stage2_kernel(const EnsDev e, const uint32_t half,
                                                    const uint32_t step, const double inv_temp,
                                                    const float* ys, const double* lsn,
                                                    const uint32_t draw, const int store,
                                                    const uint32_t slot) {
  const int hw = e.nwalk / 2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= hw) return;
  const int k = static_cast<int>(half) * hw + t;
  float* xo = e.pos + static_cast<size_t>(k) * e.npar;
  const int c = e.pass[t];
  if (c >= 0) {
    const int st = e.cst[c];
    const double lnew = e.cll[c];
    if ((st != 0 && st != kBadLike) || isnan(lnew) || isinf(lnew)) {
      e.nerr[k] += 1;
    } else {
      const double a2 = exp(inv_temp * ((lnew - e.ll[k]) - (lsn[t] - e.llsur[k])));
      if ((a2 >= 1.0) || (draws(e, step, half, k, draw).u0 < a2)) {
        const float* y = ys + static_cast<size_t>(t) * e.npar;
        for (int i = 0; i < e.npar; ++i) xo[i] = y[i];
        fill_bonus(e, xo, c);
        e.ll[k] = lnew;
        e.llsur[k] = lsn[t];
        e.naccept[k] += 1;
      }
    }
  }
  if (store) {
    float* ch = e.chain + (static_cast<size_t>(slot) * e.nwalk + k) * e.npar;
    for (int i = 0; i < e.npar; ++i) ch[i] = xo[i];
    e.chainll[static_cast<size_t>(slot) * e.nwalk + k] = e.ll[k];
  }
}

// ---- surrogate transition (Liu, Monte Carlo Strategies in Scientific
// Computing, 2001, sec. 9.4.3): with the other half fixed, each walker of
// the half runs `nsteps` stretch moves accepted on the surrogate alone
// (z^d exp((L~new - L~old)/T)), a chain reversible with respect to
// exp(L~/T); its end y is then accepted with
// exp(((L(y) - L(x)) - (L~(y) - L~(x)))/T) (stage2_kernel), which makes
// the whole move reversible with respect to the exact posterior.  One inner
// step is the delayed-acceptance first stage.  Walkers whose inner chain
// never moved need no exact evaluation.
__global__ __launch_bounds__(64) void st_init_kernel(const EnsDev e, const uint32_t half) {
  const int hw = e.nwalk / 2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= hw) return;
  const int k = static_cast<int>(half) * hw + t;
  const float* x = e.pos + static_cast<size_t>(k) * e.npar;
  float* y = e.yin + static_cast<size_t>(t) * e.npar;
  for (int i = 0; i < e.npar; ++i) y[i] = x[i];
  e.lsy[t] = e.llsur[k];
  e.moved[t] = 0;
}

// inner step of walker t: its proposal (from the inner position) against
// it on the surrogate predictions llsp; the uniform is draw `draw`
__device__ void st_inner_one(const EnsDev& e, const uint32_t half, const uint32_t step,
                             const double inv_temp, const int t, const uint32_t draw) {
  const int hw = e.nwalk / 2;
  if (!e.pvalid[t]) return;
  const int k = static_cast<int>(half) * hw + t;
  const double ls = e.llsp[t];
  if (isnan(ls) || isinf(ls)) return;
  const double a = exp(static_cast<double>(e.dexp) * log(static_cast<double>(e.z[t])) +
                       inv_temp * (ls - e.lsy[t]));
  if (!((a >= 1.0) || (draws(e, step, half, k, draw).u0 < a))) return;
  const float* p = e.prop + static_cast<size_t>(t) * e.npar;
  float* y = e.yin + static_cast<size_t>(t) * e.npar;
  for (int i = 0; i < e.npar; ++i) y[i] = p[i];
  e.lsy[t] = ls;
  e.moved[t] = 1;
}

// ---- one inner step of the surrogate transition in one launch: workgroup
// g proposes rows 16 g .. 16 g + 15 of the half from the inner positions
// `from`, predicts them with the surrogate (the network's fused inference,
// mlp_infer.h) and runs their inner acceptance.  (Delayed acceptance keeps
// separate launches: its first stage needs every workgroup's predictions,
// and a device-wide handoff inside one launch costs more than the launch
// boundary.)
__global__ __launch_bounds__(pofd_mlp_dev::kInfThreads) void inner_step_kernel(
    const EnsDev e, const uint32_t half, const uint32_t step, const float* from,
    const uint32_t draw0, const pofd_mlp_dev::InferArgs a, const double inv_temp,
    const uint32_t adraw) {
  constexpr int R = pofd_mlp_dev::kInfRows;
  const int hw = e.nwalk / 2;
  const int tid = threadIdx.x;
  const int r0 = blockIdx.x * R;
  if (tid < R && r0 + tid < hw) propose_one(e, half, step, r0 + tid, from, draw0);
  __syncthreads();  // the rows, written to global memory, visible to the workgroup
  pofd_mlp_dev::infer_rows(
      a, r0, hw, [&](int r, int c) { return e.prop[static_cast<size_t>(r) * e.npar + c]; },
      e.llsp);
  __syncthreads();
  if (tid < R && r0 + tid < hw) st_inner_one(e, half, step, inv_temp, r0 + tid, adraw);
}

__global__ __launch_bounds__(64) void bonus_init_kernel(const EnsDev e) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < e.nwalk) fill_bonus(e, e.pos + static_cast<size_t>(k) * e.npar, k);
}

// z draws for tests of generateZ's distribution
__global__ __launch_bounds__(256) void zsample_kernel(const EnsDev e, const uint32_t n,
                                                      double* out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint4 r = philox(t, 0xFFFFu, 0x2A2Au, 0x0BADF00Du, static_cast<uint32_t>(e.seed),
                         static_cast<uint32_t>(e.seed >> 32));
  const double a = e.scalefac, v = (a - 1.0) * u53(r.x, r.y) + 1.0;
  out[t] = v * v / a;
}

int launch_propose(const EnsDev& e, uint32_t half, uint32_t step, hipStream_t st,
                   const float* from, uint32_t draw0) {
  const int hw = e.nwalk / 2;
  hipLaunchKernelGGL(propose_kernel, dim3((hw + 63) / 64), dim3(64), 0, st, e, half, step, from,
                     draw0);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_gauss(const EnsDev& e, const float* x, int n, double* ll, int32_t* stv,
                 hipStream_t st) {
  hipLaunchKernelGGL(gauss_target_kernel, dim3((n + 63) / 64), dim3(64), 0, st, e, x, n, ll,
                     stv);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_accept(const EnsDev& e, uint32_t half, uint32_t step, double inv_temp, int store,
                  uint32_t slot, hipStream_t st) {
  const int hw = e.nwalk / 2;
  hipLaunchKernelGGL(accept_kernel, dim3((hw + 63) / 64), dim3(64), 0, st, e, half, step,
                     inv_temp, store, slot);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_zsample(const EnsDev& e, uint32_t n, double* out, hipStream_t st) {
  hipLaunchKernelGGL(zsample_kernel, dim3((n + 255) / 256), dim3(256), 0, st, e, n, out);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_stage1(const EnsDev& e, uint32_t half, uint32_t step, double inv_temp,
                  hipStream_t st, int moved_only, const float* ys) {
  hipLaunchKernelGGL(stage1_kernel, dim3(1), dim3(kCompactThreads), 0, st, e, half, step,
                     inv_temp, moved_only, ys ? ys : e.prop);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_stage2(const EnsDev& e, uint32_t half, uint32_t step, double inv_temp, int store,
                  uint32_t slot, hipStream_t st, const float* ys, const double* lsn,
                  uint32_t draw) {
  const int hw = e.nwalk / 2;
  hipLaunchKernelGGL(stage2_kernel, dim3((hw + 63) / 64), dim3(64), 0, st, e, half, step,
                     inv_temp, ys ? ys : e.prop, lsn ? lsn : e.llsp, draw, store, slot);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_inner_step(const EnsDev& e, uint32_t half, uint32_t step, const float* from,
                      uint32_t draw0, const pofd_mlp_dev::InferArgs& a, double inv_temp,
                      uint32_t adraw, hipStream_t st) {
  const int hw = e.nwalk / 2;
  const int R = pofd_mlp_dev::kInfRows;
  hipLaunchKernelGGL(inner_step_kernel, dim3((hw + R - 1) / R), dim3(pofd_mlp_dev::kInfThreads),
                     pofd_mlp_dev::infer_lds_bytes(a.nin, a.H), st, e, half, step, from, draw0,
                     a, inv_temp, adraw);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_st_init(const EnsDev& e, uint32_t half, hipStream_t st) {
  const int hw = e.nwalk / 2;
  hipLaunchKernelGGL(st_init_kernel, dim3((hw + 63) / 64), dim3(64), 0, st, e, half);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_bonus_init(const EnsDev& e, hipStream_t st) {
  hipLaunchKernelGGL(bonus_init_kernel, dim3((e.nwalk + 63) / 64), dim3(64), 0, st, e);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}



}  // namespace pofd_ens_dev
