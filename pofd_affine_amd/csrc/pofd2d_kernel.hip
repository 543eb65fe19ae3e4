// 2D (two-band) P(D) log-likelihood kernels for gfx950: the work of
// calcLikeDoubleSingle::getLogLike (src/calcLikeDouble.cc:558-620) for a
// batch of walkers, one beam set at a time.  Grid layout n x n, row index i
// along band 1 (RFlux1), column index j along band 2 (RFlux2), as rvals in
// PDFactoryDouble.
//
//   prep2d        model set-up per walker: knot spline, colour sigma and
//                 offset interpolants (numberCountsDoubleLogNormal::setParams,
//                 src/numberCountsDoubleLogNormal.cc:1010-1052)
//   rgrid2d       R on a grid of (f1, f2) fluxes, one row per workgroup
//                 (getR array version, :1788-2007); used for the main grid
//                 and the three edge grids of initR (src/PDFactoryDouble.cc:
//                 376-551)
//   edge2d_*      edge integrals R[0,0], R[0,y], R[x,0] in the reference's
//                 summation order, patched into the main grid
//   rstats2d      mean and variance along each axis (getRStats, :593-722)
//   rowfft2d      r2c transform of every row, columns k2 <= K2 kept
//   transpose2d   [rows][cols] <-> [cols][rows] through LDS tiles
//   colfft2d      complex transform down every kept column; spectrum
//                 exp(R^ - R^0) with the mean shifts (initPD, :1228-1340)
//   noisecol2d    per dataset: noise factor (getPD, :1396-1431) and the
//                 inverse transform down every kept column
//   rowifft2d     c2r transform of every row, clamp, row marginal
//   colsum2d      column marginal (unwrapAndNormalizePD, :1001-1011)
//   split2d       findSplitPoint on both marginals (:784-975)
//   norm2d_*      rotation by the split points, integral and means
//                 (PDDouble::getIntegral / getMeans, src/PDDouble.cc:142-452)
//   like2d        binned or unbinned bilinear log likelihood
//                 (src/PDDouble.cc:976-1222)
//   combine2d     (curr - like_offset)/like_norm over datasets
//
// Spectral pruning: after the noise factor only components with
// sigma1^2 w1^2/2 <= cut and sigma2^2 w2^2/2 <= cut survive, so the forward
// column transforms run only for k2 <= K2 and the inverse only takes k1 in
// [0, K1] U [n-K1, n) and k2 <= K2 (K1, K2 from the widest dataset).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal2d.h"

namespace pofd2 {

namespace {

constexpr int kT = 256;  // threads per workgroup

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double mul_rn(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ double add_rn(double a, double b) { return __dadd_rn(a, b); }
__device__ __forceinline__ double sub_rn(double a, double b) { return __dsub_rn(a, b); }

__device__ __forceinline__ double log2_or_small(double v) {
  return (v > 0.0) ? log2(v) : -30.0;  // pofd_mcmc::smalllogval
}

__device__ double block_sum(double v, double* red) {
  const int tid = threadIdx.x;
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  double s = red[0];
  for (int i = 1; i < nw; ++i) s += red[i];
  __syncthreads();
  return s;
}

__device__ double block_max(double v, double* red) {
  const int tid = threadIdx.x;
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  double s = red[0];
  for (int i = 1; i < nw; ++i) s = fmax(s, red[i]);
  __syncthreads();
  return s;
}

__device__ int block_max_int(int v, int* ired) {
  const int tid = threadIdx.x;
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  __syncthreads();
  if ((tid & 63) == 0) ired[tid >> 6] = v;
  __syncthreads();
  int s = ired[0];
  for (int i = 1; i < nw; ++i) s = max(s, ired[i]);
  __syncthreads();
  return s;
}

// (value, index) minimum, the lowest index among equal values; idx < 0 is
// an empty slot
__device__ void block_argmin(double& v, int& idx, double* red, int* ired) {
  const int tid = threadIdx.x;
  const int nw = blockDim.x >> 6;
  auto better = [](double a, int ia, double b, int ib) {
    if (ib < 0) return true;
    if (ia < 0) return false;
    return (a < b) || (a == b && ia < ib);
  };
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double ov = __shfl_xor(v, o);
    const int oi = __shfl_xor(idx, o);
    if (!better(v, idx, ov, oi)) {
      v = ov;
      idx = oi;
    }
  }
  __syncthreads();
  if ((tid & 63) == 0) {
    red[tid >> 6] = v;
    ired[tid >> 6] = idx;
  }
  __syncthreads();
  v = red[0];
  idx = ired[0];
  for (int i = 1; i < nw; ++i)
    if (!better(v, idx, red[i], ired[i])) {
      v = red[i];
      idx = ired[i];
    }
  __syncthreads();
}

// ---- per-walker model block (doubles): [0] status, [1] sigmult1,
// [2] sigmult2, [3] K1, [4] K2, then the knot spline (4 per interval, log2
// space), sigma and offset interpolants (4 per interval), raw sigma and
// offset values
constexpr int kMH = 8;
constexpr int kOffK = kMH, kOffS = kMH + 4 * kMaxK, kOffO = kMH + 8 * kMaxK;
constexpr int kOffSV = kMH + 12 * kMaxK, kOffOV = kMH + 13 * kMaxK;
__device__ __forceinline__ double* model_ptr(const Work2& wk, int w) {
  return wk.model + static_cast<size_t>(w) * kModelStride;
}

// largest index k with sigfac*(iflux*k)^2 >= -cut (the noise factor of the
// component is above exp(-cut)); n/2 when sigma is not positive
__device__ int kmax_axis(double sigma, double iflux, double cut, int kcap) {
  if (!(sigma > 0)) return kcap;
  const double sigfac = -0.5 * sigma * sigma;
  const double est = sqrt(2.0 * cut) / (sigma * iflux);
  long k = (est > static_cast<double>(kcap)) ? kcap : static_cast<long>(est);
  auto keep = [&](long kk) {
    const double w = mul_rn(iflux, static_cast<double>(kk));
    return mul_rn(mul_rn(sigfac, w), w) >= -cut;
  };
  while (k < kcap && keep(k + 1)) ++k;
  while (k > 0 && !keep(k)) --k;
  return static_cast<int>(k);
}

// GSL natural cubic spline coefficients (cspline_init, coeff_calc) of
// (x, y) into cf[4*i..]: y_i, b_i, c_i, d_i; no contraction, thread 0 only
__device__ void cspline_coefs(int n, const double* x, const double* y, double* cf,
                              double* scratch /* 6*kMaxK */) {
  double* c = scratch;
  double* diag = scratch + kMaxK;
  double* off = scratch + 2 * kMaxK;
  double* g = scratch + 3 * kMaxK;
  double* alpha = scratch + 4 * kMaxK;
  double* gamma = scratch + 5 * kMaxK;
  const int sys = n - 2;
  for (int i = 0; i < n; ++i) c[i] = 0.0;
  for (int i = 0; i < sys; ++i) {
    const double h_i = sub_rn(x[i + 1], x[i]);
    const double h_ip1 = sub_rn(x[i + 2], x[i + 1]);
    const double yd_i = sub_rn(y[i + 1], y[i]);
    const double yd_ip1 = sub_rn(y[i + 2], y[i + 1]);
    const double g_i = (h_i != 0.0) ? 1.0 / h_i : 0.0;
    const double g_ip1 = (h_ip1 != 0.0) ? 1.0 / h_ip1 : 0.0;
    off[i] = h_ip1;
    diag[i] = mul_rn(2.0, add_rn(h_ip1, h_i));
    g[i] = mul_rn(3.0, sub_rn(mul_rn(yd_ip1, g_ip1), mul_rn(yd_i, g_i)));
  }
  if (sys == 1) {
    c[1] = g[0] / diag[0];
  } else if (sys > 1) {
    double* z = diag;
    alpha[0] = diag[0];
    gamma[0] = off[0] / alpha[0];
    for (int i = 1; i < sys - 1; ++i) {
      alpha[i] = sub_rn(diag[i], mul_rn(off[i - 1], gamma[i - 1]));
      gamma[i] = off[i] / alpha[i];
    }
    alpha[sys - 1] = sub_rn(diag[sys - 1], mul_rn(off[sys - 2], gamma[sys - 2]));
    z[0] = g[0];
    for (int i = 1; i < sys; ++i) z[i] = sub_rn(g[i], mul_rn(gamma[i - 1], z[i - 1]));
    for (int i = 0; i < sys; ++i) z[i] = z[i] / alpha[i];
    c[sys] = z[sys - 1];
    for (int i = sys - 2; i >= 0; --i) c[i + 1] = sub_rn(z[i], mul_rn(gamma[i], c[i + 2]));
  }
  for (int i = 0; i < n - 1; ++i) {
    const double dx = sub_rn(x[i + 1], x[i]);
    const double dy = sub_rn(y[i + 1], y[i]);
    cf[4 * i] = y[i];
    cf[4 * i + 1] = sub_rn(dy / dx, mul_rn(dx, add_rn(c[i + 1], mul_rn(2.0, c[i]))) / 3.0);
    cf[4 * i + 2] = c[i];
    cf[4 * i + 3] = sub_rn(c[i + 1], c[i]) / mul_rn(3.0, dx);
  }
}

// gsl_interp_bsearch over x[0..n-1]
__device__ __forceinline__ int bsearch_k(const double* x, int n, double v) {
  int lo = 0, hi = n - 1;
  while (hi > lo + 1) {
    const int m = (lo + hi) >> 1;
    if (x[m] > v)
      hi = m;
    else
      lo = m;
  }
  return lo;
}

__device__ __forceinline__ double spline_eval(const double* x, const double* cf, int n,
                                              double v) {
  const int k = bsearch_k(x, n, v);
  const double dx = sub_rn(v, x[k]);
  const double* c = cf + 4 * k;
  return add_rn(c[0], mul_rn(dx, add_rn(c[1], mul_rn(dx, add_rn(c[2], mul_rn(dx, c[3]))))));
}

// getSigmaInner / getOffsetInner (src/numberCountsDoubleLogNormal.cc:1172-1193)
__device__ double colour_eval(const double* x, const double* vals, const double* cf,
                              int n, double f1, double floorv, bool floored) {
  if (n == 1) return vals[0];
  if (f1 <= x[0]) return vals[0];
  if (f1 >= x[n - 1]) return vals[n - 1];
  double val;
  if (n == 2) {  // gsl_interp_linear
    const double dx = sub_rn(x[1], x[0]);
    val = (dx > 0.0) ? add_rn(vals[0], mul_rn(sub_rn(f1, x[0]) / dx, sub_rn(vals[1], vals[0])))
                     : 0.0;
  } else {
    val = spline_eval(x, cf, n, f1);
  }
  if (floored && val < floorv) val = floorv;
  return val;
}

}  // namespace

// ===========================================================================
// prep2d: model set-up, one thread per walker
// ===========================================================================
__global__ __launch_bounds__(64) void prep2d_kernel(const Set2Dev s, const Work2 wk,
                                                    const float* __restrict__ theta,
                                                    const uint32_t stride,
                                                    const int32_t* __restrict__ prev_status,
                                                    const int accumulate) {
  __shared__ double scratch[6 * kMaxK];
  __shared__ double vals[kMaxK];
  if (threadIdx.x != 0) return;
  const int w = blockIdx.x;
  double* m = model_ptr(wk, w);
  const float* th = theta + static_cast<size_t>(w) * stride;
  const int nk = s.nk, ns = s.ns, no = s.no;
  int status = accumulate ? prev_status[w] : POFD_OK;
  const double sm1 = th[nk + ns + no], sm2 = th[nk + ns + no + 1];
  // setParams and the validity checks (src/numberCountsDoubleLogNormal.cc:
  // 1010-1119): log2 knot values, sigma and offset values as doubles
  bool valid = true;
  for (int i = 0; i < nk; ++i) {
    vals[i] = mul_rn(s.logfac, static_cast<double>(th[i]));
    if (isnan(vals[i])) valid = false;
  }
  if (valid) cspline_coefs(nk, s.logknots, vals, m + kOffK, scratch);
  for (int i = 0; i < ns; ++i) {
    const double v = th[nk + i];
    if (isnan(v) || isinf(v)) valid = false;
    vals[i] = v;
    m[kOffSV + i] = v;
  }
  if (valid && ns > 2) cspline_coefs(ns, s.sigmaknots, vals, m + kOffS, scratch);
  for (int i = 0; i < no; ++i) {
    const double v = th[nk + ns + i];
    if (isnan(v)) valid = false;
    vals[i] = v;
    m[kOffOV + i] = v;
  }
  if (valid && no > 2) cspline_coefs(no, s.offsetknots, vals, m + kOffO, scratch);
  if (status == POFD_OK) {
    if (isnan(sm1) || isnan(sm2))
      status = POFD_E_ARG;
    else if (mul_rn(s.maxsig1, sm1) < 0.0 || mul_rn(s.maxsig2, sm2) < 0.0)
      status = POFD_E_BADLIKE;  // calcLikeDoubleSingle::bad_like, :571-574
    else if (!valid || !(s.min_sigma >= 0.0))
      status = POFD_E_INVALID_MODEL;
  }
  // spectral bounds of the widest dataset
  int K1 = 0, K2 = 0;
  for (int f = 0; f < s.nfields; ++f) {
    K1 = max(K1, kmax_axis(sm1 * s.fields[f].sigma_base1, s.iflux1, s.spectral_cut, s.n / 2));
    K2 = max(K2, kmax_axis(sm2 * s.fields[f].sigma_base2, s.iflux2, s.spectral_cut, s.n / 2));
  }
  m[0] = status;
  m[1] = sm1;
  m[2] = sm2;
  m[3] = K1;
  m[4] = K2;
  m[5] = valid ? 1.0 : 0.0;
}

// ===========================================================================
// bonus2d: getNS, getFluxPerArea and getFluxSqPerArea of both bands
// (numberCountsDoubleLogNormal::splineInt, src/numberCountsDoubleLogNormal.cc:
// 1549-1673, integrand evalPowfNDoubleLogNormal :2424-2519):
// int S1^(a+b) exp(b mu(S1) + b^2/2 sigma(S1)^2) n1(S1) dS1 as Gauss-Legendre
// sums in log2 S1 between consecutive knots of the band-1, sigma and offset
// models (analytic there).  One 64-lane workgroup per walker.
// ===========================================================================
__global__ __launch_bounds__(64) void bonus2d_kernel(const Set2Dev s, const Work2 wk) {
  __shared__ double red[1];
  const int w = blockIdx.x;
  const double* m = model_ptr(wk, w);
  double* bo = wk.bonus + static_cast<size_t>(w) * 8;
  if (m[5] == 0.0) {  // invalid model: splineInt returns NaN
    if (threadIdx.x < 5) bo[threadIdx.x] = __longlong_as_double(0x7ff8000000000000ULL);
    return;
  }
  const int nk = s.nk, ns = s.ns, no = s.no;
  const double* kcf = m + kOffK;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const int npts = (s.nbp - 1) * 32;
  for (int t = threadIdx.x; t < npts; t += 64) {
    const int k = t >> 5, g = t & 31;
    const double v0 = s.bpts[k], v1 = s.bpts[k + 1];
    if (!(v1 > v0)) continue;
    const double h = 0.5 * (v1 - v0), c = 0.5 * (v1 + v0);
    const double v = c + h * s.glx[g];
    const double s1 = exp2(v);
    const double lc = spline_eval(s.logknots, kcf, nk, v);
    double sg = colour_eval(s.sigmaknots, m + kOffSV, m + kOffS, ns, s1, s.min_sigma, true);
    if (sg < s.min_sigma) sg = s.min_sigma;
    const double mu = colour_eval(s.offsetknots, m + kOffOV, m + kOffO, no, s1, 0.0, false);
    const double base = s.glw[g] * h * exp2(lc + v);  // S1 n1(S1) per unit log2 S1
    const double e1 = exp(mu + 0.5 * sg * sg);         // beta = 1
    const double e2 = exp(2.0 * mu + 2.0 * sg * sg);   // beta = 2
    acc[0] += base;
    acc[1] += base * s1;
    acc[2] += base * s1 * e1;
    acc[3] += base * s1 * s1;
    acc[4] += base * s1 * s1 * e2;
  }
  const double ln2 = 0.69314718055994530942;
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    const double tot = block_sum(acc[q], red);
    if (threadIdx.x == 0) bo[q] = ln2 * tot;
  }
}

// ===========================================================================
// rgrid2d: R on (f1[i], f2[j]) for every walker; one workgroup per
// (row block, walker).  out[i*ostride + j] = R * scale.
// ===========================================================================
// The pieces of R that depend on the row flux f1 but not on f2, per row
// of the grid and beam element (RWork, :1868-1950): weight (0 where the
// element adds nothing), log-flux centre and Gaussian width factor, into
// wk.rpieces for rgrid2d_kernel.  One thread per (row, element), so that
// their logs, exp2s and colour splines run in parallel rather than on a
// few threads of each row's workgroup.
__global__ __launch_bounds__(kT) void rpieces2d_kernel(const Set2Dev s, const Work2 wk,
                                                       const RGrid g) {
  __shared__ double lk[kMaxK], lkcf[4 * kMaxK];
  __shared__ double sk[kMaxK], scf[4 * kMaxK], svl[kMaxK];
  __shared__ double ok[kMaxK], ocf[4 * kMaxK], ovl[kMaxK];
  const int w = blockIdx.y;
  const double* m = model_ptr(wk, w);
  if (m[0] != POFD_OK) return;  // uniform
  const int nk = s.nk, ns = s.ns, no = s.no;
  for (int t = threadIdx.x; t < kMaxK; t += kT) {
    lk[t] = s.logknots[t];
    sk[t] = s.sigmaknots[t];
    ok[t] = s.offsetknots[t];
    svl[t] = m[kOffSV + t];
    ovl[t] = m[kOffOV + t];
  }
  for (int t = threadIdx.x; t < 4 * kMaxK; t += kT) {
    lkcf[t] = m[kOffK + t];
    scf[t] = m[kOffS + t];
    ocf[t] = m[kOffO + t];
  }
  __syncthreads();
  const int maxc = wk.rpieces_w / 3;
  const int t = blockIdx.x * kT + threadIdx.x;
  const int i = t / maxc, e = t - i * maxc;
  if (i >= g.n1) return;
  double f1val = g.f1[i];
  if ((f1val == 0) || (f1val < 0 && !s.hasneg1)) return;  // a row of zeros
  int sgn1 = 0;
  if (!(f1val > 0)) {
    sgn1 = 2;
    f1val = fabs(f1val);
  }
  const int np = s.comp[sgn1].n, nn = s.comp[sgn1 + 1].n;
  if (e >= np + nn) return;
  const double minknot = s.knots[0], maxknot = s.knots[nk - 1];
  const double ilog2toe = 1.4426950408889634074;  // 1/ln 2
  const double logf1 = log(f1val);
  const int k = (e < np) ? 0 : 1;
  const BeamComp& bc = s.comp[sgn1 + k];
  const int jj = (k == 0) ? e : e - np;
  double* c = wk.rpieces + static_cast<size_t>(w) * wk.rpieces_stride +
              static_cast<size_t>(i) * wk.rpieces_w + 3 * e;
  double c0 = 0.0, c1 = 0.0, c2 = 0.0;
  if (f1val < bc.maxf1) {
#pragma clang fp contract(off)
    const double ieta1 = bc.ieta1[jj];
    const double f1prod = f1val * ieta1;
    if (f1prod >= minknot && f1prod < maxknot) {
      const double cts = exp2(spline_eval(lk, lkcf, nk, log2(f1prod)));
      if (cts > 0) {
        const double isigma = 1.0 / colour_eval(sk, svl, scf, ns, f1prod, s.min_sigma, true);
        c0 = (bc.w ? bc.w[jj] * ieta1 * isigma * cts * s.prefac
                   : cts * ieta1 * isigma * s.prefac);
        c1 = logf1 + colour_eval(ok, ovl, ocf, no, f1prod, 0.0, false) - bc.logratio[jj];
        c2 = -0.5 * isigma * isigma * ilog2toe;
      }
    }
  }
  c[0] = c0;
  c[1] = c1;
  c[2] = c2;
}

// R on one flux grid, a row per workgroup step: the row's pieces (from
// rpieces2d_kernel) into LDS, then every f2 of the row sums its beam
// elements' Gaussian colour terms in element order.
__global__ __launch_bounds__(kT) void rgrid2d_kernel(const Set2Dev s, const Work2 wk,
                                                     const RGrid g) {
  extern __shared__ double rw[];  // 3 * (np + nn) per row
  const int w = blockIdx.y;
  const double* m = model_ptr(wk, w);
  if (m[0] != POFD_OK) return;  // uniform
  const double* pieces = wk.rpieces + static_cast<size_t>(w) * wk.rpieces_stride;
  for (int i = blockIdx.x; i < g.n1; i += gridDim.x) {
    double f1val = g.f1[i];
    if ((f1val == 0) || (f1val < 0 && !s.hasneg1)) {
      for (int j = threadIdx.x; j < g.n2; j += kT)
        g.out[static_cast<size_t>(w) * g.wstride + static_cast<size_t>(i) * g.ostride + j] = 0.0;
      continue;
    }
    const int sgn1 = (f1val > 0) ? 0 : 2;
    const int np = s.comp[sgn1].n, nn = s.comp[sgn1 + 1].n;
    const double* prow = pieces + static_cast<size_t>(i) * wk.rpieces_w;
    for (int e = threadIdx.x; e < 3 * (np + nn); e += kT) rw[e] = prow[e];
    __syncthreads();
    double* orow = g.out + static_cast<size_t>(w) * g.wstride + static_cast<size_t>(i) * g.ostride;
    for (int j = threadIdx.x; j < g.n2; j += kT) {
      const double if2 = g.f2inv[j];
      double r = 0.0;
      if (if2 != 0) {
        const int fs = g.f2sgn[j];
        const int sgn = sgn1 + fs;
        if (s.comp[sgn].n > 0) {
#pragma clang fp contract(off)
          const int cn = s.comp[sgn].n;
          const double* base = rw + 3 * (fs ? np : 0);
          const double lf2 = g.f2log[j];
          double workval = 0.0;
          for (int k = 0; k < cn; ++k) {
            const double c0 = base[3 * k];
            if (c0 != 0) {
              const double tfac = lf2 - base[3 * k + 1];
              workval += c0 * exp2(tfac * tfac * base[3 * k + 2]);
            }
          }
          r = (workval * if2) * g.scale;
        }
      }
      orow[j] = r;
    }
    __syncthreads();
  }
}

// ===========================================================================
// edge integrals (initR with setEdge, src/PDFactoryDouble.cc:386-544)
// ===========================================================================
// Ry[i] = trapezoid over the linear edge in band 2 of R(RFlux1[i], edge2)
// and the y integral of the nedge x nedge grid (row sums), thread per row
__global__ void edge2d_rows_kernel(const Set2Dev s, const Work2 wk, const int nrows,
                                   const double* __restrict__ grid, const size_t gstride,
                                   double* __restrict__ out, const size_t ostride,
                                   const double norm) {
  const int w = blockIdx.y;
  if (model_ptr(wk, w)[0] != POFD_OK) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  const int ne = s.nedge;
  const double* r = grid + static_cast<size_t>(w) * gstride + static_cast<size_t>(i) * ne;
  double sc;
  {
#pragma clang fp contract(off)
    sc = 0.5 * r[0];
    for (int j = 1; j < ne - 1; ++j) sc += r[j];
    sc += 0.5 * r[ne - 1];
    if (norm != 0.0) sc *= norm;
  }
  out[static_cast<size_t>(w) * ostride + i] = sc;
}

// Rx[j] = log-x trapezoid over the nedge rows of R(edge1, RFlux2[j])
__global__ void edge2d_cols_kernel(const Set2Dev s, const Work2 wk) {
  const int w = blockIdx.y;
  if (model_ptr(wk, w)[0] != POFD_OK) return;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = s.n, ne = s.nedge;
  if (j >= n) return;
  const double* ex = wk.ex + static_cast<size_t>(w) * wk.ex_stride;
  double acc;
  {
#pragma clang fp contract(off)
    acc = 0.5 * s.edge1[0] * ex[j];
    for (int i = 1; i < ne - 1; ++i) acc += s.edge1[i] * ex[static_cast<size_t>(i) * n + j];
    acc += 0.5 * s.edge1[ne - 1] * ex[static_cast<size_t>(ne - 1) * n + j];
    acc *= s.iRxnorm;
  }
  wk.rx[static_cast<size_t>(w) * wk.edge_stride + j] = acc;
}

// R[0,0] from the row sums of the nedge x nedge grid, then the patch of row
// 0 and column 0 of the main grid (times dflux1*dflux2)
__global__ void edge2d_patch_kernel(const Set2Dev s, const Work2 wk) {
  const int w = blockIdx.y;
  if (model_ptr(wk, w)[0] != POFD_OK) return;
  const int n = s.n, ne = s.nedge;
  double* R = wk.R + static_cast<size_t>(w) * wk.R_stride;
  const double* rx = wk.rx + static_cast<size_t>(w) * wk.edge_stride;
  const double* ry = wk.ry + static_cast<size_t>(w) * wk.edge_stride;
  const double* r00rows = wk.r00rows + static_cast<size_t>(w) * wk.edge_stride;
  const double fluxfac = s.dflux1 * s.dflux2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) {
#pragma clang fp contract(off)
    double sc = 0.5 * s.edge1[0] * r00rows[0];
    for (int i = 1; i < ne - 1; ++i) sc += s.edge1[i] * r00rows[i];
    sc += 0.5 * s.edge1[ne - 1] * r00rows[ne - 1];
    R[0] = (sc * s.iR00norm) * fluxfac;
  }
  if (t >= 1 && t < n) {
#pragma clang fp contract(off)
    R[t] = rx[t] * fluxfac;
    R[static_cast<size_t>(t) * n] = (ry[t] * s.iRynorm) * fluxfac;
  }
}

// ===========================================================================
// rstats2d: trapezoid moments of R along each axis; partials per block row
// ===========================================================================
__global__ __launch_bounds__(kT) void rstats2d_kernel(const Set2Dev s, const Work2 wk) {
  __shared__ double red[kT / 64];
  const int w = blockIdx.y;
  if (model_ptr(wk, w)[0] != POFD_OK) return;
  const int n = s.n;
  const double* R = wk.R + static_cast<size_t>(w) * wk.R_stride;
  double m1 = 0, m2 = 0, v1 = 0, v2 = 0;
  const int rows_per = (n + gridDim.x - 1) / gridDim.x;
  const int i0 = blockIdx.x * rows_per, i1 = min(n, i0 + rows_per);
  for (int i = i0; i < i1; ++i) {
    const double wi = (i == 0 || i == n - 1) ? 0.5 : 1.0;
    const double cf1 = s.rflux1[i];
    for (int j = threadIdx.x; j < n; j += kT) {
      const double wj = (j == 0 || j == n - 1) ? 0.5 : 1.0;
      const double cR = wi * R[static_cast<size_t>(i) * n + j];
      const double p1 = wj * cf1 * cR;
      m1 += p1;
      v1 += cf1 * p1;
      const double cf2 = s.rflux2[j];
      const double p2 = wj * cf2 * cR;
      m2 += p2;
      v2 += cf2 * p2;
    }
  }
  double* part = wk.rstat_part + (static_cast<size_t>(w) * gridDim.x + blockIdx.x) * 4;
  const double a = block_sum(m1, red), b = block_sum(m2, red);
  const double c = block_sum(v1, red), d = block_sum(v2, red);
  if (threadIdx.x == 0) {
    part[0] = a;
    part[1] = b;
    part[2] = c;
    part[3] = d;
  }
}

// ===========================================================================
// LDS transforms: complex length L (power of two, <= 4096), radix-4 Stockham
// plus a final radix-2 stage when log2 L is odd.  tw[k] = e^{-2 pi i k/L}.
// SIGN -1 forward, +1 unnormalised backward.  Returns the buffer holding the
// result.
// ===========================================================================
template <int SIGN>
__device__ double2* fft_lds(double2* a, double2* b, int L, const double2* __restrict__ tw) {
  int Ns = 1;
  double2* src = a;
  double2* dst = b;
  while (Ns * 4 <= L) {
    const int q = L >> 2;
    const int tstride = L / (4 * Ns);
    for (int j = threadIdx.x; j < q; j += blockDim.x) {
      double2 v0 = src[j], v1 = src[j + q], v2 = src[j + 2 * q], v3 = src[j + 3 * q];
      const int jm = j & (Ns - 1);
      if (Ns > 1) {
        double2 t1 = tw[jm * tstride], t2 = tw[2 * jm * tstride], t3 = tw[3 * jm * tstride];
        if (SIGN > 0) {
          t1.y = -t1.y;
          t2.y = -t2.y;
          t3.y = -t3.y;
        }
        v1 = cmul(v1, t1);
        v2 = cmul(v2, t2);
        v3 = cmul(v3, t3);
      }
      const double2 a0 = cadd(v0, v2), a1 = csub(v0, v2);
      const double2 a2 = cadd(v1, v3), a3 = csub(v1, v3);
      double2 y1, y3;
      if (SIGN < 0) {
        y1 = make_double2(a1.x + a3.y, a1.y - a3.x);
        y3 = make_double2(a1.x - a3.y, a1.y + a3.x);
      } else {
        y1 = make_double2(a1.x - a3.y, a1.y + a3.x);
        y3 = make_double2(a1.x + a3.y, a1.y - a3.x);
      }
      const int d = (j / Ns) * Ns * 4 + jm;
      dst[d] = cadd(a0, a2);
      dst[d + Ns] = y1;
      dst[d + 2 * Ns] = csub(a0, a2);
      dst[d + 3 * Ns] = y3;
    }
    __syncthreads();
    double2* t = src;
    src = dst;
    dst = t;
    Ns *= 4;
  }
  if (Ns < L) {  // radix-2 stage, Ns = L/2
    const int h = L >> 1;
    for (int j = threadIdx.x; j < h; j += blockDim.x) {
      double2 v0 = src[j], v1 = src[j + h];
      double2 t1 = tw[j];
      if (SIGN > 0) t1.y = -t1.y;
      v1 = cmul(v1, t1);
      dst[j] = cadd(v0, v1);
      dst[j + h] = csub(v0, v1);
    }
    __syncthreads();
    src = dst;
  }
  return src;
}

// ===========================================================================
// rowfft2d: r2c of every row of R (length n), outputs k2 = 0..K2 into
// Y[i][k2] (row stride ystride complex).  One workgroup per (row, walker).
// ===========================================================================
__global__ __launch_bounds__(kT) void rowfft2d_kernel(const Set2Dev s, const Work2 wk) {
  extern __shared__ double2 lds[];
  const int w = blockIdx.y;
  const double* m = model_ptr(wk, w);
  if (m[0] != POFD_OK) return;
  const int K2 = static_cast<int>(m[4]);
  const int n = s.n, h = n / 2;
  double2* a = lds;
  double2* b = lds + h;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const double2* row =
        reinterpret_cast<const double2*>(wk.R + static_cast<size_t>(w) * wk.R_stride +
                                         static_cast<size_t>(i) * n);
    for (int m = threadIdx.x; m < h; m += kT) a[m] = row[m];  // x[2m] + i x[2m+1]
    __syncthreads();
    const double2* z = fft_lds<-1>(a, b, h, s.tw_half);
    double2* Y = wk.Y + static_cast<size_t>(w) * wk.Y_stride + static_cast<size_t>(i) * wk.ystride;
    for (int k = threadIdx.x; k <= K2; k += kT) {
      const int kk = (k == h) ? 0 : k, mk = (k == 0) ? 0 : h - k;
      const double2 zk = z[kk];
      const double2 zm = make_double2(z[mk].x, -z[mk].y);
      const double2 A = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y + zm.y));
      const double2 B = make_double2(0.5 * (zk.y - zm.y), -0.5 * (zk.x - zm.x));
      const double2 wv = (k == h) ? make_double2(-1.0, 0.0) : s.tw_n[k];
      Y[k] = cadd(A, cmul(B, wv));
    }
    __syncthreads();
  }
}

// ===========================================================================
// transpose2d: out[c][r] = in[r][c] for r < nr, c < nc (complex), 32x32
// LDS tiles; grid (tiles, walkers)
// ===========================================================================
// kmode 1: the column count of `in` is K2+1 of the walker (Y -> YT);
// kmode 2: the row count of `in` is K2+1 (Q -> QT)
__global__ __launch_bounds__(kT) void transpose2d_kernel(const Work2 wk, const double2* in,
                                                         size_t in_wstride, int in_stride,
                                                         double2* out, size_t out_wstride,
                                                         int out_stride, int nr, int nc,
                                                         int kmode) {
  __shared__ double2 tile[32][33];
  const int w = blockIdx.z;
  const double* m = model_ptr(wk, w);
  if (m[0] != POFD_OK) return;
  const int kk = static_cast<int>(m[4]) + 1;
  if (kmode == 1) nc = kk;
  if (kmode == 2) nr = kk;
  const int tc = blockIdx.x * 32, tr = blockIdx.y * 32;
  if (tc >= nc || tr >= nr) return;
  const int lx = threadIdx.x & 31, ly = threadIdx.x >> 5;  // 32 x 8
  const double2* src = in + static_cast<size_t>(w) * in_wstride;
  double2* dst = out + static_cast<size_t>(w) * out_wstride;
  for (int y = ly; y < 32; y += 8) {
    const int r = tr + y, c = tc + lx;
    tile[y][lx] = (r < nr && c < nc) ? src[static_cast<size_t>(r) * in_stride + c]
                                     : make_double2(0.0, 0.0);
  }
  __syncthreads();
  for (int y = ly; y < 32; y += 8) {
    const int c = tc + y, r = tr + lx;
    if (c < nc && r < nr) dst[static_cast<size_t>(c) * out_stride + r] = tile[lx][y];
  }
}

// ===========================================================================
// colfft2d: forward complex transform (length n) of every kept column
// YT[k2][:], then the spectrum exp(X - r0) with the shifts into P[k2][k1]
// (only kept k1).  One workgroup per (k2, walker).
// ===========================================================================
__global__ __launch_bounds__(kT) void colfft2d_kernel(const Set2Dev s, const Work2 wk) {
  extern __shared__ double2 lds[];
  const int w = blockIdx.y;
  const double* m = model_ptr(wk, w);
  if (m[0] != POFD_OK) return;
  const int K1 = static_cast<int>(m[3]), K2 = static_cast<int>(m[4]);
  const int n = s.n;
  double2* a = lds;
  double2* b = lds + n;
  const double* st = wk.stats + static_cast<size_t>(w) * kStats;
  const double r0 = st[ST2_R0];
  const bool ds1 = st[ST2_DOSHIFT1] != 0.0, ds2 = st[ST2_DOSHIFT2] != 0.0;
  const double sf1 = st[ST2_SHIFT1] * s.iflux1, sf2 = st[ST2_SHIFT2] * s.iflux2;
  const int ncplx = n / 2 + 1;
  for (int k2 = blockIdx.x; k2 <= K2; k2 += gridDim.x) {
    const double2* col = wk.YT + static_cast<size_t>(w) * wk.YT_stride + static_cast<size_t>(k2) * n;
    for (int t = threadIdx.x; t < n; t += kT) a[t] = col[t];
    __syncthreads();
    const double2* X = fft_lds<-1>(a, b, n, s.tw_n);
    double2* pr = wk.PR + static_cast<size_t>(w) * wk.PR_stride + static_cast<size_t>(k2) * n;
    for (int k1 = threadIdx.x; k1 < n; k1 += kT) {
      const bool kept = (k1 <= K1) || (k1 >= n - K1);
      double2 v = make_double2(0.0, 0.0);
      if (kept) {
#pragma clang fp contract(off)
        const double2 xv = X[k1];
        const double rval = xv.x - r0;
        double ival = xv.y;
        double sp1 = 0.0;
        if (ds1) sp1 = (k1 < ncplx) ? sf1 * static_cast<double>(k1) : -sf1 * static_cast<double>(n - k1);
        if (ds1 && ds2)
          ival = ival - sp1 - sf2 * static_cast<double>(k2);
        else if (ds1)
          ival = ival - sp1;
        else if (ds2)
          ival = ival - sf2 * static_cast<double>(k2);
        const double ef = exp(rval);
        double sn, cs;
        sincos(ival, &sn, &cs);
        v = make_double2(ef * cs, ef * sn);
      }
      if (k1 == 0 && k2 == 0) v = make_double2(1.0, 0.0);
      pr[k1] = v;
    }
    __syncthreads();
  }
}

// ===========================================================================
// noisecol2d: per dataset, noise factor times the spectrum and the inverse
// complex transform down every kept column into Q[k2][i]
// ===========================================================================
__global__ __launch_bounds__(kT) void noisecol2d_kernel(const Set2Dev s, const Work2 wk,
                                                        const int f) {
  extern __shared__ double2 lds[];
  const int w = blockIdx.y;
  const double* m = model_ptr(wk, w);
  if (m[0] != POFD_OK) return;
  const int K1 = static_cast<int>(m[3]), K2 = static_cast<int>(m[4]);
  const int n = s.n, ncplx = n / 2 + 1;
  double2* a = lds;
  double2* b = lds + n;
  const double sigma1 = m[1] * s.fields[f].sigma_base1;
  const double sigma2 = m[2] * s.fields[f].sigma_base2;
  const bool noisy = (sigma1 > 0 || sigma2 > 0);
  const double sigfac1 = -0.5 * sigma1 * sigma1 * s.iflux1 * s.iflux1;
  const double sigfac2 = -0.5 * sigma2 * sigma2 * s.iflux2 * s.iflux2;
  for (int k2 = blockIdx.x; k2 <= K2; k2 += gridDim.x) {
    const double2* pr = wk.PR + static_cast<size_t>(w) * wk.PR_stride + static_cast<size_t>(k2) * n;
    for (int k1 = threadIdx.x; k1 < n; k1 += kT) {
      double2 v = pr[k1];
      const bool kept = (k1 <= K1) || (k1 >= n - K1);
      if (!kept) {
        v = make_double2(0.0, 0.0);
      } else if (noisy) {
#pragma clang fp contract(off)
        const double didx1 = (k1 < ncplx) ? static_cast<double>(k1) : static_cast<double>(n - k1);
        const double didx2 = static_cast<double>(k2);
        const double ef = exp(sigfac1 * didx1 * didx1 + sigfac2 * didx2 * didx2);
        v = make_double2(ef * v.x, ef * v.y);
      }
      if (k1 == 0 && k2 == 0) v = make_double2(1.0, 0.0);
      a[k1] = v;
    }
    __syncthreads();
    const double2* q = fft_lds<+1>(a, b, n, s.tw_n);
    double2* Q = wk.Q + static_cast<size_t>(w) * wk.Q_stride + static_cast<size_t>(k2) * n;
    for (int i = threadIdx.x; i < n; i += kT) Q[i] = q[i];
    __syncthreads();
  }
}

// ===========================================================================
// rowifft2d: c2r of every row from QT[i][0..n/2] (zero beyond K2), clamp,
// store P and the row trapezoid marginal
// ===========================================================================
__global__ __launch_bounds__(kT) void rowifft2d_kernel(const Set2Dev s, const Work2 wk,
                                                       const int f) {
  extern __shared__ double2 lds[];
  __shared__ double red[kT / 64];
  const int w = blockIdx.y;
  const double* m = model_ptr(wk, w);
  if (m[0] != POFD_OK) return;
  const int K2 = static_cast<int>(m[4]);
  const int n = s.n, h = n / 2;
  double2* a = lds;
  double2* b = lds + h;
  const size_t wf = static_cast<size_t>(w) * s.nfields + f;
  constexpr int kMaxPairs = 4096 / 2 / kT;  // output pairs per thread, n <= 4096
  double cacc[2 * kMaxPairs];
#pragma unroll
  for (int q = 0; q < 2 * kMaxPairs; ++q) cacc[q] = 0.0;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const double2* X = wk.QT + static_cast<size_t>(w) * wk.QT_stride + static_cast<size_t>(i) * wk.ystride;
    for (int k = threadIdx.x; k < h; k += kT) {
      double2 xa = (k <= K2) ? X[k] : make_double2(0.0, 0.0);
      const int mk = h - k;
      double2 xb = (mk <= K2) ? X[mk] : make_double2(0.0, 0.0);
      if (k == 0) {  // FFTW's c2r ignores Im of the k2 = 0 and n/2 entries
        xa.y = 0.0;
        xb.y = 0.0;
      }
      const double2 E = make_double2(xa.x + xb.x, xa.y - xb.y);
      const double2 D = make_double2(xa.x - xb.x, xa.y + xb.y);
      const double2 wv = make_double2(s.tw_n[k].x, -s.tw_n[k].y);  // e^{+2 pi i k/n}
      const double2 O = cmul(D, wv);
      a[k] = make_double2(E.x - O.y, E.y + O.x);
    }
    __syncthreads();
    const double2* z = fft_lds<+1>(a, b, h, s.tw_half);
    double* P = wk.P + wf * wk.P_stride + static_cast<size_t>(i) * n;
    const double wi = (i == 0 || i == n - 1) ? 0.5 : 1.0;
    double part = 0.0;
#pragma unroll
    for (int q = 0; q < kMaxPairs; ++q) {
      const int m2 = threadIdx.x + q * kT;
      if (m2 < h) {
        double x0 = z[m2].x, x1 = z[m2].y;
        if (x0 < 0) x0 = 0.0;  // unwrapAndNormalizePD positivity
        if (x1 < 0) x1 = 0.0;
        P[2 * m2] = x0;
        P[2 * m2 + 1] = x1;
        const double w0 = (2 * m2 == 0) ? 0.5 : 1.0;
        const double w1 = (2 * m2 + 1 == n - 1) ? 0.5 : 1.0;
        part += w0 * x0 + w1 * x1;
        cacc[2 * q] += wi * x0;
        cacc[2 * q + 1] += wi * x1;
      }
    }
    const double rs = block_sum(part, red);
    if (threadIdx.x == 0) wk.rsum1[wf * wk.edge_stride + i] = rs;
  }
  // this workgroup's share of the column trapezoid sums
  double* cp = wk.colpart + static_cast<size_t>(w) * wk.colpart_stride +
               static_cast<size_t>(blockIdx.x) * n;
#pragma unroll
  for (int q = 0; q < kMaxPairs; ++q) {
    const int m2 = threadIdx.x + q * kT;
    if (m2 < h) {
      cp[2 * m2] = cacc[2 * q];
      cp[2 * m2 + 1] = cacc[2 * q + 1];
    }
  }
}

// column trapezoid marginal (unwrapAndNormalizePD, src/PDFactoryDouble.cc:
// 1001-1011) from the per-workgroup partial sums of rowifft2d, thread per
// column
__global__ void colsum2d_kernel(const Set2Dev s, const Work2 wk, const int f) {
  const int w = blockIdx.y;
  if (model_ptr(wk, w)[0] != POFD_OK) return;
  const int n = s.n;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const size_t wf = static_cast<size_t>(w) * s.nfields + f;
  const double* cp = wk.colpart + static_cast<size_t>(w) * wk.colpart_stride + j;
  double acc = 0.0;
  for (int b = 0; b < kRowBlocks2; ++b) acc += cp[static_cast<size_t>(b) * n];
  wk.rsum2[wf * wk.edge_stride + j] = acc;
}

// ===========================================================================
// split2d: findSplitPoint (src/PDFactoryDouble.cc:784-975) on the marginal of
// one axis; one workgroup per (axis, walker)
// ===========================================================================
__global__ __launch_bounds__(kT) void split2d_kernel(const Set2Dev s, const Work2 wk,
                                                     const int f) {
  __shared__ double red[kT / 64];
  __shared__ int ired[kT / 64];
  __shared__ int ires[3];
  __shared__ double dres[1];
  const int axis = blockIdx.x, w = blockIdx.y;
  const double* m = model_ptr(wk, w);
  if (m[0] != POFD_OK) return;  // uniform
  const int tid = threadIdx.x;
  const size_t wf = static_cast<size_t>(w) * s.nfields + f;
  double* dbg = wk.dbg + wf * kDbg2;
  const double* st = wk.stats + static_cast<size_t>(w) * kStats;
  const double sigma = m[1 + axis] * (axis ? s.fields[f].sigma_base2 : s.fields[f].sigma_base1);
  const double var = st[ST2_VAR1 + axis];
  const double sgpos = sqrt(var + sigma * sigma);
  const double sgneg = (9 * sigma * sigma > var) ? sigma : sgpos;  // getPD, :1413-1420
  const double* data = (axis ? wk.rsum2 : wk.rsum1) + wf * wk.edge_stride;
  const double dflux = axis ? s.dflux2 : s.dflux1;
  const int cs = s.n;
  const double n1 = 2.25, f1 = 5e-14, f2 = 1.5e-5, f3 = 5;
  int status = POFD_OK;
  if (sgpos <= 0 || sgneg <= 0) status = POFD_E_SPLIT;
  const double fluxrange = cs * dflux;
  const double n1sigpos = n1 * sgpos, n1signeg = n1 * sgneg;
  if (fluxrange < (n1sigpos + n1signeg)) status = POFD_E_SPLIT;
  auto clampi = [](double v) { return static_cast<int>(v > 2e9 ? 2e9 : (v < -2e9 ? -2e9 : v)); };
  const int topidx = cs - clampi(n1signeg / dflux) - 1;
  const int botidx = clampi(n1sigpos / dflux);
  if (topidx >= cs || botidx < 0 || botidx > topidx) status = POFD_E_SPLIT;
  int split = -1;
  if (status == POFD_OK) {  // uniform across the workgroup
    double pk = -1.0;
    for (int i = tid; i < cs; i += kT) pk = fmax(pk, data[i]);
    const double peak = block_max(pk, red);
    // first minimum over [botidx, topidx]
    double mv = 0.0;
    int mi = -1;
    for (int i = botidx + tid; i <= topidx; i += kT) {
      const double v = data[i];
      if (mi < 0 || v < mv) {
        mv = v;
        mi = i;
      }
    }
    block_argmin(mv, mi, red, ired);
    if (tid == 0) {
      int stt = (peak <= 0.0 || mv > f2 * peak) ? POFD_E_SPLIT : POFD_OK;
      int decided = 0, sp = -1;
      double minval = mv, jitter = 0.0;
      const int minidx = mi;
      if (stt == POFD_OK) {
        const int pre_jitter_n = 17;
        if (minidx > pre_jitter_n) {
          bool mono = true;
          const int initidx = minidx - pre_jitter_n;
          double prev = data[initidx];
          jitter = prev;
          for (int i = initidx + 1; i < minidx; ++i) {
            const double cv = data[i];
            if (cv >= prev) mono = false;
            jitter += cv;
            prev = cv;
          }
          if (mono) {
            sp = minidx;
            decided = 1;
          } else {
            jitter += minval;
            jitter = 1.1 * jitter / static_cast<double>(pre_jitter_n + 1);
          }
        } else {
          jitter = f1 * peak;
          if (minval > jitter) {
            sp = minidx;
            decided = 1;
          }
        }
      }
      if (stt == POFD_OK && !decided) {
        if (minval <= 0.0) {
          const int nminup = 3;
          int bi, ti;
          if (minidx < botidx)
            bi = botidx;
          else
            bi = (minidx >= nminup) ? (minidx - nminup) : minidx;
          ti = (minidx > topidx - nminup) ? topidx - nminup : minidx + nminup;
          ti = (ti >= cs) ? (cs - 1) : ti;
          for (int i = bi; i < ti; ++i)
            if (data[i] > minval) minval = data[i];
        }
        if (minval > 0) jitter = (f3 * minval) < jitter ? (f3 * minval) : jitter;
        if (botidx >= topidx || data[topidx] < jitter) stt = POFD_E_SPLIT;
      }
      ires[0] = stt;
      ires[1] = decided;
      ires[2] = sp;
      dres[0] = jitter;
    }
    __syncthreads();
    status = ires[0];
    split = ires[2];
    if (status == POFD_OK && !ires[1]) {
      // the largest i in [botidx, topidx - 1] with data[i] < jitter
      const double jitter = dres[0];
      int cand = -1;
      for (int i = botidx + tid; i <= topidx - 1; i += kT)
        if (data[i] < jitter) cand = i;
      const int found = block_max_int(cand, ired);
      if (found < 0)
        status = POFD_E_SPLIT;
      else
        split = found + 1;
    }
  }
  if (tid == 0) {
    dbg[D2_ST1 + axis] = status;
    dbg[D2_SPLIT1 + axis] = split;
    dbg[D2_SG1POS + axis] = sgpos;
    dbg[D2_SIGMA1 + axis] = sigma;
  }
}

// norm2d_rows: per rotated row i' the trapezoid row sum A and the weighted
// sum B = sum_j' j' P (last weight halved), block per (row, walker)
__global__ __launch_bounds__(kT) void norm2d_rows_kernel(const Set2Dev s, const Work2 wk,
                                                         const int f) {
  __shared__ double red[kT / 64];
  const int w = blockIdx.y;
  if (model_ptr(wk, w)[0] != POFD_OK) return;
  const size_t wf = static_cast<size_t>(w) * s.nfields + f;
  const double* dbg = wk.dbg + wf * kDbg2;
  if (dbg[D2_ST1] != POFD_OK || dbg[D2_ST2] != POFD_OK) return;
  const int n = s.n;
  const int s1 = static_cast<int>(dbg[D2_SPLIT1]), s2 = static_cast<int>(dbg[D2_SPLIT2]);
  const double* P = wk.P + wf * wk.P_stride;
  for (int ip = blockIdx.x; ip < n; ip += gridDim.x) {
    const int i = (ip + s1) % n;
    const double* row = P + static_cast<size_t>(i) * n;
    double a = 0.0, bsum = 0.0;
    for (int jp = threadIdx.x; jp < n; jp += kT) {
      const double v = row[(jp + s2) % n];
      const double wj = (jp == 0 || jp == n - 1) ? 0.5 : 1.0;
      a += wj * v;
      if (jp >= 1) bsum += (jp == n - 1 ? 0.5 : 1.0) * static_cast<double>(jp) * v;
    }
    const double A = block_sum(a, red);
    const double B = block_sum(bsum, red);
    if (threadIdx.x == 0) {
      wk.rowA[wf * wk.edge_stride + ip] = A;
      wk.rowB[wf * wk.edge_stride + ip] = B;
    }
  }
}

// norm2d_final: the split status, then integral, means and minflux from
// the rotated row sums (PDDouble::getIntegral / getMeans), one workgroup
// per walker
__global__ __launch_bounds__(kT) void norm2d_final_kernel(const Set2Dev s, const Work2 wk,
                                                          const int f) {
  __shared__ double red[kT / 64];
  const int w = blockIdx.x;
  if (model_ptr(wk, w)[0] != POFD_OK) return;
  const size_t wf = static_cast<size_t>(w) * s.nfields + f;
  double* dbg = wk.dbg + wf * kDbg2;
  const int st1 = static_cast<int>(dbg[D2_ST1]), st2 = static_cast<int>(dbg[D2_ST2]);
  const int sst = (st1 != POFD_OK) ? st1 : st2;  // band 1 is searched first
  if (sst != POFD_OK) {
    if (threadIdx.x == 0) dbg[D2_STATUS] = sst;
    return;
  }
  const int n = s.n;
  const double* A = wk.rowA + wf * wk.edge_stride;
  const double* B = wk.rowB + wf * wk.edge_stride;
  double tp = 0.0, m1p = 0.0, m2p = 0.0;
  for (int i = threadIdx.x; i < n; i += kT) {
    const double wi = (i == 0 || i == n - 1) ? 0.5 : 1.0;
    tp += wi * A[i];
    m1p += wi * static_cast<double>(i) * A[i];
    m2p += wi * B[i];
  }
  const double tot = block_sum(tp, red);
  const double mean1 = block_sum(m1p, red);
  const double mean2 = block_sum(m2p, red);
  if (threadIdx.x == 0) {
    const double integral = tot * s.dflux1 * s.dflux2;
    const double itot = 1.0 / integral;
    // getMeans on the normalised array (PDDouble.cc:355-452, donorm false)
    const double m1 = mean1 * itot * (s.dflux1 * s.dflux1 * s.dflux2);
    const double m2 = mean2 * itot * (s.dflux1 * s.dflux2 * s.dflux2);
    int status = POFD_OK;
    if (isinf(m1) || isnan(m1) || isinf(m2) || isnan(m2)) status = POFD_E_MEAN;
    dbg[D2_STATUS] = status;
    dbg[D2_INTEGRAL] = integral;
    dbg[D2_ITOT] = itot;
    dbg[D2_MINFLUX1] = -m1;
    dbg[D2_MINFLUX2] = -m2;
  }
}

// like2d: bilinear interpolation of log2 of the normalised, rotated P(D)
// over bins (binned) or pixel pairs; partial sums per block
__global__ __launch_bounds__(kT) void like2d_kernel(const Set2Dev s, const Work2 wk,
                                                    const int f) {
  __shared__ double red[kT / 64];
  const int w = blockIdx.y;
  const size_t wf = static_cast<size_t>(w) * s.nfields + f;
  double* part = wk.lpart + wf * kLikeBlocks + blockIdx.x;
  const double* dbg = wk.dbg + wf * kDbg2;
  if (model_ptr(wk, w)[0] != POFD_OK || dbg[D2_STATUS] != POFD_OK) {
    if (threadIdx.x == 0) *part = 0.0;
    return;
  }
  const Field2Dev& fd = s.fields[f];
  const int n = s.n;
  const int s1 = static_cast<int>(dbg[D2_SPLIT1]), s2 = static_cast<int>(dbg[D2_SPLIT2]);
  const double itot = dbg[D2_ITOT];
  const double minflux1 = dbg[D2_MINFLUX1], minflux2 = dbg[D2_MINFLUX2];
  const double dflux1 = s.dflux1, dflux2 = s.dflux2;
  const double* P = wk.P + wf * wk.P_stride;
  auto L = [&](int i, int j) {  // log2 of the rotated, normalised P(D)
    int ii = i + s1;
    if (ii >= n) ii -= n;
    int jj = j + s2;
    if (jj >= n) jj -= n;
    return log2_or_small(P[static_cast<size_t>(ii) * n + jj] * itot);
  };
  double acc = 0.0;
  {
#pragma clang fp contract(off)
    const double maxflux1 = minflux1 + static_cast<double>(n - 1) * dflux1;
    const double maxflux2 = minflux2 + static_cast<double>(n - 1) * dflux2;
    const double idflux1 = 1.0 / dflux1, idflux2 = 1.0 / dflux2;
    const uint64_t nitems = fd.binned ? static_cast<uint64_t>(fd.nb1) * fd.nb2 : fd.npix;
    const uint64_t per = (nitems + kLikeBlocks - 1) / kLikeBlocks;
    const uint64_t i0 = per * blockIdx.x;
    const uint64_t i1 = (i0 + per < nitems) ? i0 + per : nitems;
    for (uint64_t it = i0 + threadIdx.x; it < i1; it += kT) {
      double c1, c2, wgt;
      if (fd.binned) {
        const uint32_t nin = fd.bins[it];
        if (nin == 0) continue;
        const uint32_t b1 = static_cast<uint32_t>(it / fd.nb2), b2 = static_cast<uint32_t>(it % fd.nb2);
        c1 = fd.bincent01 + static_cast<double>(b1) * fd.bindelta1;
        c2 = fd.bincent02 + static_cast<double>(b2) * fd.bindelta2;
        wgt = static_cast<double>(nin);
      } else {
        c1 = fd.d1[it];
        c2 = fd.d2[it];
        wgt = 1.0;
      }
      const double delt1 = (c1 - minflux1) * idflux1;
      const double delt2 = (c2 - minflux2) * idflux2;
      const int idx1 = static_cast<int>(delt1), idx2 = static_cast<int>(delt2);
      double iv;
      // binned uses </> at the edges, unbinned <=/>= (PDDouble.cc:1029-1048,
      // 1140-1158)
      const bool lo1 = fd.binned ? (c1 < minflux1) : (c1 <= minflux1);
      const bool hi1 = fd.binned ? (c1 > maxflux1) : (c1 >= maxflux1);
      const bool lo2 = fd.binned ? (c2 < minflux2) : (c2 <= minflux2);
      const bool hi2 = fd.binned ? (c2 > maxflux2) : (c2 >= maxflux2);
      if (lo1) {
        iv = lo2 ? L(0, 0) : (hi2 ? L(0, n - 1) : L(0, idx2));
      } else if (hi1) {
        iv = lo2 ? L(n - 1, 0) : (hi2 ? L(n - 1, n - 1) : L(n - 1, idx2));
      } else if (lo2) {
        iv = L(idx1, 0);
      } else if (hi2) {
        iv = L(idx1, n - 1);
      } else {
        const double t = delt1 - static_cast<double>(idx1);
        const double u = delt2 - static_cast<double>(idx2);
        const double omu = 1.0 - u, omt = 1.0 - t;
        iv = omt * (omu * L(idx1, idx2) + u * L(idx1, idx2 + 1)) +
             t * (omu * L(idx1 + 1, idx2) + u * L(idx1 + 1, idx2 + 1));
      }
      acc += wgt * iv;
    }
  }
  const double sb = block_sum(acc, red);
  if (threadIdx.x == 0) *part = sb;
}

// combine2d: per walker, datasets in order
__global__ void combine2d_kernel(const Set2Dev s, const Work2 wk, const int accumulate,
                                 double* __restrict__ loglike, int32_t* __restrict__ status) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= wk.nwalk) return;
  const double* m = model_ptr(wk, w);
  const int st = static_cast<int>(m[0]);
  const double prev = accumulate ? loglike[w] : 0.0;
  if (st == POFD_E_BADLIKE) {
    status[w] = st;
    loglike[w] = prev + 1e25;
    return;
  }
  if (st != POFD_OK) {
    status[w] = st;
    loglike[w] = __longlong_as_double(0x7ff8000000000000ULL);
    return;
  }
  double total = 0.0;
  for (int f = 0; f < s.nfields; ++f) {
    const size_t wf = static_cast<size_t>(w) * s.nfields + f;
    double* dbg = wk.dbg + wf * kDbg2;
    const int fs = static_cast<int>(dbg[D2_STATUS]);
    if (fs != POFD_OK) {
      status[w] = fs;
      loglike[w] = __longlong_as_double(0x7ff8000000000000ULL);
      return;
    }
    const double* lp = wk.lpart + wf * kLikeBlocks;
    double s2 = 0.0;
    for (int b = 0; b < kLikeBlocks; ++b) s2 += lp[b];
    const double curr = 0.69314718055994530942 * s2;  // log2toe
    dbg[D2_LL] = curr;
    total += (curr - s.fields[f].like_offset) / s.fields[f].like_norm;
  }
  status[w] = POFD_OK;
  loglike[w] = prev + total;
}

// combine the rstats partials, decide the shifts, record r0 (after rowfft/
// colfft of column k2 = 0 this is done by colfft's input: r0 = sum R)
__global__ void rstats2d_final_kernel(const Set2Dev s, const Work2 wk, const int nparts) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= wk.nwalk) return;
  if (model_ptr(wk, w)[0] != POFD_OK) return;
  const double* part = wk.rstat_part + static_cast<size_t>(w) * nparts * 4;
  double m1 = 0, m2 = 0, v1 = 0, v2 = 0;
  for (int b = 0; b < nparts; ++b) {
    m1 += part[4 * b];
    m2 += part[4 * b + 1];
    v1 += part[4 * b + 2];
    v2 += part[4 * b + 3];
  }
  double* st = wk.stats + static_cast<size_t>(w) * kStats;
  st[ST2_MN1] = m1;
  st[ST2_MN2] = m2;
  st[ST2_VAR1] = v1;
  st[ST2_VAR2] = v2;
  const double sg1 = sqrt(v1), sg2 = sqrt(v2);
  const bool d1 = (sg1 > s.dflux1) && (fabs(m1) > s.dflux1);
  const bool d2 = (sg2 > s.dflux2) && (fabs(m2) > s.dflux2);
  st[ST2_DOSHIFT1] = d1 ? 1.0 : 0.0;
  st[ST2_DOSHIFT2] = d2 ? 1.0 : 0.0;
  st[ST2_SHIFT1] = d1 ? -m1 : 0.0;
  st[ST2_SHIFT2] = d2 ? -m2 : 0.0;
}

// r0 = Re X[0][0]: the k2 = 0 column transform at k1 = 0 is the sum of the
// k2 = 0 row outputs; computed by one block per walker before colfft2d
__global__ __launch_bounds__(kT) void r0_2d_kernel(const Set2Dev s, const Work2 wk) {
  __shared__ double red[kT / 64];
  const int w = blockIdx.x;
  if (model_ptr(wk, w)[0] != POFD_OK) return;
  const double2* col = wk.YT + static_cast<size_t>(w) * wk.YT_stride;  // k2 = 0
  double a = 0.0;
  for (int i = threadIdx.x; i < s.n; i += kT) a += col[i].x;
  const double r0 = block_sum(a, red);
  if (threadIdx.x == 0) wk.stats[static_cast<size_t>(w) * kStats + ST2_R0] = r0;
}

}  // namespace pofd2

namespace pofd2 {

int launch_loglike2d(const Set2Dev& s, const Work2& wk, const RGrid& gmain,
                     const RGrid& g00, const RGrid& gx, const RGrid& gy,
                     const float* d_theta, uint32_t stride, double* d_ll,
                     int32_t* d_st, int accumulate, void* stream, hipEvent_t* ev) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int W = wk.nwalk, n = s.n, ne = s.nedge;
  auto mark = [&](int i) {
    if (ev) (void)hipEventRecord(ev[i], st);
  };
  // rgrid2d: dynamic LDS for the per-row beam pieces of two sign components
  int maxc = 0;
  for (int k = 0; k < 2; ++k) maxc = std::max(maxc, s.comp[2 * k].n + s.comp[2 * k + 1].n);
  const size_t rg_lds = static_cast<size_t>(3) * std::max(maxc, 1) * sizeof(double);
  if (rg_lds > kRgridMaxLds) return POFD_E_ARG;
  if (rg_lds + 8192 > 65536)  // + the static knot tables (4.3 KB)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(rgrid2d_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(rg_lds));
  const int rblocks = 64;  // row blocks per walker (grid-stride over rows)
  // one R grid: its per-row beam pieces, then the grid
  auto rgrid = [&](const RGrid& g, const int blocks) {
    const int pb = static_cast<int>((static_cast<size_t>(g.n1) * (wk.rpieces_w / 3) + kT - 1) / kT);
    hipLaunchKernelGGL(rpieces2d_kernel, dim3(pb, W), dim3(kT), 0, st, s, wk, g);
    hipLaunchKernelGGL(rgrid2d_kernel, dim3(blocks, W), dim3(kT), rg_lds, st, s, wk, g);
  };
  mark(POFD_STAGE_PREP);
  hipLaunchKernelGGL(prep2d_kernel, dim3(W), dim3(64), 0, st, s, wk, d_theta, stride, d_st,
                     accumulate);
  if (s.bonus) hipLaunchKernelGGL(bonus2d_kernel, dim3(W), dim3(64), 0, st, s, wk);
  mark(POFD_STAGE_FORWARD);
  if (s.edge_integ) {
    rgrid(g00, std::min(rblocks, ne));
    rgrid(gx, std::min(rblocks, ne));
    rgrid(gy, rblocks);
    hipLaunchKernelGGL(edge2d_rows_kernel, dim3((ne + 255) / 256, W), dim3(256), 0, st, s, wk,
                       ne, wk.e00, g00.wstride, wk.r00rows, wk.edge_stride, 0.0);
    hipLaunchKernelGGL(edge2d_rows_kernel, dim3((n + 255) / 256, W), dim3(256), 0, st, s, wk,
                       n, wk.ey, gy.wstride, wk.ry, wk.edge_stride, 0.0);
    hipLaunchKernelGGL(edge2d_cols_kernel, dim3((n + 255) / 256, W), dim3(256), 0, st, s, wk);
  }
  rgrid(gmain, rblocks * 4);
  if (s.edge_integ)
    hipLaunchKernelGGL(edge2d_patch_kernel, dim3((n + 255) / 256, W), dim3(256), 0, st, s, wk);
  const int nparts = 64;
  hipLaunchKernelGGL(rstats2d_kernel, dim3(nparts, W), dim3(kT), 0, st, s, wk);
  hipLaunchKernelGGL(rstats2d_final_kernel, dim3((W + 63) / 64), dim3(64), 0, st, s, wk, nparts);
  const size_t lds_row = static_cast<size_t>(n) * sizeof(double2);      // 2 x n/2
  const size_t lds_col = static_cast<size_t>(2 * n) * sizeof(double2);  // 2 x n
  if (lds_row + 1024 > 65536) {  // n = 4096: 64 KB plus the static reduction slots
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(rowfft2d_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(lds_row));
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(rowifft2d_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(lds_row));
  }
  if (lds_col > 65536) {  // n = 4096: 128 KB of the CU's 160 KB
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(colfft2d_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(lds_col));
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(noisecol2d_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(lds_col));
  }
  hipLaunchKernelGGL(rowfft2d_kernel, dim3(256, W), dim3(kT), lds_row, st, s, wk);
  const int nc = n / 2 + 1;
  hipLaunchKernelGGL(transpose2d_kernel, dim3((nc + 31) / 32, (n + 31) / 32, W), dim3(kT), 0, st,
                     wk, static_cast<const double2*>(wk.Y), wk.Y_stride, wk.ystride, wk.YT,
                     wk.YT_stride, n, n, nc, 1);
  hipLaunchKernelGGL(r0_2d_kernel, dim3(W), dim3(kT), 0, st, s, wk);
  hipLaunchKernelGGL(colfft2d_kernel, dim3(64, W), dim3(kT), lds_col, st, s, wk);
  mark(POFD_STAGE_INVERSE);
  for (int f = 0; f < s.nfields; ++f) {
    hipLaunchKernelGGL(noisecol2d_kernel, dim3(64, W), dim3(kT), lds_col, st, s, wk, f);
    hipLaunchKernelGGL(transpose2d_kernel, dim3((n + 31) / 32, (nc + 31) / 32, W), dim3(kT), 0,
                       st, wk, static_cast<const double2*>(wk.Q), wk.Q_stride, n, wk.QT,
                       wk.QT_stride, wk.ystride, nc, n, 2);
    hipLaunchKernelGGL(rowifft2d_kernel, dim3(kRowBlocks2, W), dim3(kT), lds_row, st, s, wk, f);
    hipLaunchKernelGGL(colsum2d_kernel, dim3((n + 255) / 256, W), dim3(256), 0, st, s, wk, f);
  }
  mark(POFD_STAGE_FINALIZE);
  for (int f = 0; f < s.nfields; ++f) {
    hipLaunchKernelGGL(split2d_kernel, dim3(2, W), dim3(kT), 0, st, s, wk, f);
    hipLaunchKernelGGL(norm2d_rows_kernel, dim3(256, W), dim3(kT), 0, st, s, wk, f);
    hipLaunchKernelGGL(norm2d_final_kernel, dim3(W), dim3(kT), 0, st, s, wk, f);
  }
  mark(POFD_STAGE_PIXELS);
  for (int f = 0; f < s.nfields; ++f)
    hipLaunchKernelGGL(like2d_kernel, dim3(kLikeBlocks, W), dim3(kT), 0, st, s, wk, f);
  mark(POFD_STAGE_COMBINE);
  hipLaunchKernelGGL(combine2d_kernel, dim3((W + 63) / 64), dim3(64), 0, st, s, wk, accumulate,
                     d_ll, d_st);
  mark(POFD_NSTAGES);
  return hipGetLastError() == hipSuccess ? POFD_OK : POFD_E_INTERNAL;
}

// calcLikeDouble::getLogLike's prior terms (src/calcLikeDouble.cc:
// 1230-1280): Gaussian priors on the two sigma multipliers (mean 1), on the
// mean flux per area of each band (the CFIRB) and on the mean flux^2 per
// area of each band (the Poisson noise), and the difference regularisation
// of the band-1 knot slopes (numberCountsDoubleLogNormal::
// differenceRegularize, src/numberCountsDoubleLogNormal.cc:2154-2173)
__global__ void prior2d_kernel(const Set2Dev s, const Work2 w, const Prior2Dev pr,
                               const float* __restrict__ theta, const uint32_t stride,
                               const uint32_t nwalk, double* __restrict__ loglike,
                               const int32_t* __restrict__ status) {
  const uint32_t walker = blockIdx.x * blockDim.x + threadIdx.x;
  if (walker >= nwalk) return;
  const int st = status[walker];
  if (st != POFD_OK && st != POFD_E_BADLIKE) return;
  const double h = 0.918938517570495605469;  // log(2 pi) / 2
  const float* t = theta + static_cast<size_t>(walker) * stride;
  const double* bo = w.bonus + static_cast<size_t>(walker) * 8;
  const int nmp = s.nk + s.ns + s.no;
  double add = 0.0;
  for (int b = 0; b < 2; ++b) {
    if (pr.flags & (kP2Sigma1 << b)) {
      const double v = (static_cast<double>(t[nmp + b]) - 1.0) / pr.sigma_width[b];
      add -= h + log(pr.sigma_width[b]) + 0.5 * v * v;
    }
  }
  for (int b = 0; b < 2; ++b) {
    if (pr.flags & (kP2Cfirb1 << b)) {
      const double v = (pr.cfirb_mean[b] - bo[1 + b]) / pr.cfirb_sigma[b];
      add -= h + log(pr.cfirb_sigma[b]) + 0.5 * v * v;
    }
  }
  for (int b = 0; b < 2; ++b) {
    if (pr.flags & (kP2Poisson1 << b)) {
      const double v = (pr.poisson_mean[b] - bo[3 + b]) / pr.poisson_sigma[b];
      add -= h + log(pr.poisson_sigma[b]) + 0.5 * v * v;
    }
  }
  if (pr.reg_alpha > 0.0 && s.nk > 1) {
    const int n = s.nk;
    double mean = 0.0;
    for (int i = 1; i < n; ++i)
      mean += (s.logfac * static_cast<double>(t[i]) - s.logfac * static_cast<double>(t[i - 1])) /
              (s.logknots[i] - s.logknots[i - 1]);
    mean /= static_cast<double>(n - 1);
    double pen = 0.0;
    for (int i = 1; i < n; ++i) {
      const double d =
          (s.logfac * static_cast<double>(t[i]) - s.logfac * static_cast<double>(t[i - 1])) /
              (s.logknots[i] - s.logknots[i - 1]) -
          mean;
      pen -= d * d;
    }
    add += pr.reg_alpha * pen;
  }
  loglike[walker] += add;
}

int launch_prior2d(const Set2Dev& s0, const Work2& w0, const Prior2Dev& pr, const float* d_theta,
                   uint32_t stride, uint32_t nwalk, double* d_ll, const int32_t* d_st,
                   void* stream) {
  hipLaunchKernelGGL(prior2d_kernel, dim3((nwalk + 63) / 64), dim3(64), 0,
                     static_cast<hipStream_t>(stream), s0, w0, pr, d_theta, stride, nwalk, d_ll,
                     d_st);
  return hipGetLastError() == hipSuccess ? POFD_OK : POFD_E_INTERNAL;
}

}  // namespace pofd2
