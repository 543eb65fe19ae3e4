// Fused 1D P(D) log-likelihood kernel for gfx950: one workgroup (1024
// threads) evaluates one walker's log likelihood for one beam set, i.e. the
// work of calcLikeSingle::getLogLike (src/calcLike.cc:389-443):
//
//   initPD  (src/PDFactory.cc:1031-1141)
//     initRInterp: R at the log-spaced interpolation nodes against the beam
//                  histogram (numberCountsKnotsSpline::getR,
//                  src/numberCountsKnotsSpline.cc:436-529), log2, natural
//                  cubic spline over the nodes (GSL cspline)
//     initR:       R at every flux of the transform grid (:541-620)
//     getMeanVarFromR (:625-679)
//     r2c FFT + exp(R^ - R^0) with the mean shift (:1094-1132)
//   getPD   (src/PDFactory.cc:1154-1243) for every dataset of the beam set
//     noise factor, c2r FFT, clamp, findSplitPoint (:691-921), rotate,
//     normalise, mean (:934-1009), log2 (src/PD.cc:169-184)
//   PD::getLogLike binned / unbinned (src/PD.cc:557-677)
//
// Transform design.  The real transform of length N = 2M is done as a
// complex transform of length M = 256 * M2 in the four-step form
// (m = M2*n1 + n2, k = k1 + 256*k2): 256-point column transforms in LDS, a
// twiddle, then the transform across columns.  After the noise factor only
// spectral indices k <= kmax survive (exp(-sigma^2 w^2 / 2) below the
// configured cut), so the across-column transform is evaluated only for the
// rows k2 holding k <= kmax or k >= M - kmax, as a direct sum accumulated
// while the columns stream through LDS.  The inverse mirrors it.  The R
// values are generated inside the forward column loop and never stored, so
// the only HBM traffic per walker and dataset is the P(D) itself.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace pofd {

namespace {

constexpr int kWaves = kThreads / 64;
constexpr int kPad = kM1 + 1;  // padded column stride in LDS (double2)

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // a*conj(b)
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 cconj(double2 a) {
  return make_double2(a.x, -a.y);
}

// LDS layout of the fused kernel.
struct Smem {
  double2 buf[kChunk * kPad];   // 16 columns of 256 (+1 pad); also aliased
                                // by the node-spline arrays during setup
  double2 w256[kM1];
  double2 wlo[kM1];
  double2 wm2[512];
  double kn[4][kMaxKnots];      // knot spline y, b, c, d per interval
  double ks[7][kMaxKnots];      // knot spline solve scratch (thread 0)
  double red[kWaves * 4];
  int rowk2[512];               // spectral rows (k2 values) in use
  int rowof[512];               // k2 -> row position, -1 if absent
  double sc[32];                // broadcast scalars
  int isc[32];
  unsigned long long tstamp[8]; // wall-clock phase stamps (thread 0)
};

enum {  // sc[] slots
  S_SIGMULT = 0, S_MN, S_VAR, S_R0, S_SHIFT, S_PEAK, S_MINVAL, S_SUM, S_SUMJ,
  S_JITTER, S_ITOT, S_MINFLUX, S_LL, S_SGPOS, S_SGNEG, S_INTEGRAL, S_SHIGH,
  S_TMP0, S_TMP1
};
enum {  // isc[] slots
  I_STATUS = 0, I_NROWS, I_KMAXB, I_DOSHIFT, I_MINIDX, I_BOT, I_TOP,
  I_SPLIT, I_DONE, I_FOUND, I_KMAXF, I_TMP0, I_TMP1
};

// ---- block-wide reductions (deterministic fixed trees) ---------------------
__device__ double block_sum(double v, double* red) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  double s = red[0];
#pragma unroll
  for (int i = 1; i < kWaves; ++i) s += red[i];
  return s;
}

__device__ double block_max(double v, double* red) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double u = __shfl_xor(v, o);
    v = (u > v) ? u : v;
  }
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  double s = red[0];
#pragma unroll
  for (int i = 1; i < kWaves; ++i) s = (red[i] > s) ? red[i] : s;
  return s;
}

__device__ int block_max_int(int v, double* red) {
  return static_cast<int>(block_max(static_cast<double>(v), red));
}

// (value, index) minimum with first-index tie break; idx < 0 means empty
__device__ void block_argmin(double& v, int& idx, double* red) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double u = __shfl_xor(v, o);
    int ui = __shfl_xor(idx, o);
    bool take = (ui >= 0) && (idx < 0 || u < v || (u == v && ui < idx));
    v = take ? u : v;
    idx = take ? ui : idx;
  }
  __syncthreads();
  if ((tid & 63) == 0) {
    red[tid >> 6] = v;
    red[kWaves + (tid >> 6)] = static_cast<double>(idx);
  }
  __syncthreads();
  double bv = red[0];
  int bi = static_cast<int>(red[kWaves]);
  for (int i = 1; i < kWaves; ++i) {
    double u = red[i];
    int ui = static_cast<int>(red[kWaves + i]);
    bool take = (ui >= 0) && (bi < 0 || u < bv || (u == bv && ui < bi));
    bv = take ? u : bv;
    bi = take ? ui : bi;
  }
  v = bv;
  idx = bi;
}

// ---- 256-point transforms of 16 LDS columns, radix-4 Stockham ------------
// Each column is handled by 64 threads holding 4 points.  SIGN = -1 is the
// forward (FFTW_FORWARD) direction, +1 the unnormalised backward one.
template <int SIGN>
__device__ void fft256_cols(double2* buf, const double2* w256) {
  const int tid = threadIdx.x;
  const int col = tid >> 6, j = tid & 63;
  double2* c = buf + col * kPad;
#pragma unroll
  for (int Ns = 1; Ns < kM1; Ns *= 4) {
    double2 v0 = c[j], v1 = c[j + 64], v2 = c[j + 128], v3 = c[j + 192];
    const int jm = j & (Ns - 1);
    if (Ns > 1) {
      const int e = jm * (64 / Ns);
      double2 t1 = w256[e], t2 = w256[2 * e], t3 = w256[3 * e];
      if (SIGN > 0) {
        t1 = cconj(t1);
        t2 = cconj(t2);
        t3 = cconj(t3);
      }
      v1 = cmul(v1, t1);
      v2 = cmul(v2, t2);
      v3 = cmul(v3, t3);
    }
    const double2 a0 = cadd(v0, v2), a1 = csub(v0, v2);
    const double2 a2 = cadd(v1, v3), a3 = csub(v1, v3);
    double2 y0 = cadd(a0, a2), y2 = csub(a0, a2), y1, y3;
    if (SIGN < 0) {  // a1 -/+ i a3
      y1 = make_double2(a1.x + a3.y, a1.y - a3.x);
      y3 = make_double2(a1.x - a3.y, a1.y + a3.x);
    } else {
      y1 = make_double2(a1.x - a3.y, a1.y + a3.x);
      y3 = make_double2(a1.x + a3.y, a1.y - a3.x);
    }
    __syncthreads();
    const int d = (j / Ns) * Ns * 4 + jm;
    c[d] = y0;
    c[d + Ns] = y1;
    c[d + 2 * Ns] = y2;
    c[d + 3 * Ns] = y3;
    __syncthreads();
  }
}

// The reference's index and flux arithmetic is done without contraction so
// that a*b+c rounds exactly as the reference's (non-FMA) build does.
__device__ __forceinline__ double mul_rn(double a, double b) {
  return __dmul_rn(a, b);
}
__device__ __forceinline__ double add_rn(double a, double b) {
  return __dadd_rn(a, b);
}
__device__ __forceinline__ double sub_rn(double a, double b) {
  return __dsub_rn(a, b);
}

// Largest spectral index whose noise factor survives the cut:
// sigfac*w*w >= -cut with w = iflux*k (src/PDFactory.cc:1180-1182).
__device__ int spectral_kmax(double sigma, double iflux, double cut, int M) {
  if (!(sigma > 0)) return M;
  const double sigfac = -0.5 * sigma * sigma;
  double est = sqrt(2.0 * cut) / (sigma * iflux);
  long k = (est > static_cast<double>(M)) ? M : static_cast<long>(est);
  auto keep = [&](long kk) {
    const double w = mul_rn(iflux, static_cast<double>(kk));
    return mul_rn(mul_rn(sigfac, w), w) >= -cut;
  };
  while (k < M && keep(k + 1)) ++k;
  while (k > 0 && !keep(k)) --k;
  return static_cast<int>(k);
}

__device__ __forceinline__ double log2_or_small(double v) {
  return (v > 0.0) ? log2(v) : -30.0;  // pofd_mcmc::smalllogval
}

}  // namespace

// One workgroup per walker.
__global__ __launch_bounds__(kThreads) void loglike1d_kernel(
    const BeamsetDev bs, const WorkDev wk, const int bsidx,
    const float* __restrict__ theta, const uint32_t stride,
    double* __restrict__ loglike, int32_t* __restrict__ status,
    const int accumulate) {
  __shared__ Smem sm;
  const int tid = threadIdx.x;
  const int walker = blockIdx.x;
  const int N = bs.N, M = bs.M, M2 = bs.M2;
  const int ninterp = bs.ninterp, nk = bs.nknots;
  const double dflux = bs.dflux;

  double* coef = wk.coef + walker * wk.coef_stride;
  double* ynode_g = wk.ynode + walker * wk.ynode_stride;
  double2* zrows = wk.zrows + walker * wk.zrows_stride;
  double2* prtrans = wk.prtrans + walker * wk.spec_stride;
  double2* xp = wk.xp + walker * wk.spec_stride;

  // ------------------------------------------------------------------ setup
  if (tid == 0) sm.tstamp[0] = wall_clock64();
  if (tid == 0) {
    int st = accumulate ? status[walker] : POFD_OK;
    sm.isc[I_STATUS] = st;
    const float* th = theta + static_cast<size_t>(walker) * stride;
    double* lkv = sm.ks[6];
    bool valid = true;
    for (int i = 0; i < nk; ++i) {
      // numberCountsKnotsSpline::setParams src/numberCountsKnotsSpline.cc:265-274
      lkv[i] = mul_rn(bs.logfac, static_cast<double>(th[i]));
      if (isnan(lkv[i])) valid = false;
    }
    const double sigmult = static_cast<double>(th[nk]);
    sm.sc[S_SIGMULT] = sigmult;
    if (st == POFD_OK) {
      if (isnan(sigmult)) {
        st = POFD_E_ARG;
      } else if (mul_rn(bs.maxsigma_base, sigmult) < 0.0) {
        st = POFD_E_BADLIKE;  // calcLikeSingle::bad_like, calcLike.cc:403
      } else if (!valid) {
        st = POFD_E_INVALID_MODEL;  // PDFactory.cc:1046-1047
      }
      sm.isc[I_STATUS] = st;
    }
    if (st == POFD_OK) {
      // GSL natural cubic spline through (logknots, logknotvals):
      // cspline_init + gsl_linalg_solve_symm_tridiag
      const double* xa = bs.logknots;
      const int sys = nk - 2;
      double* c = sm.ks[0];
      double* diag = sm.ks[1];
      double* off = sm.ks[2];
      double* g = sm.ks[3];
      for (int i = 0; i < nk; ++i) c[i] = 0.0;
      for (int i = 0; i < sys; ++i) {
        const double h_i = sub_rn(xa[i + 1], xa[i]);
        const double h_ip1 = sub_rn(xa[i + 2], xa[i + 1]);
        const double yd_i = sub_rn(lkv[i + 1], lkv[i]);
        const double yd_ip1 = sub_rn(lkv[i + 2], lkv[i + 1]);
        const double g_i = (h_i != 0.0) ? 1.0 / h_i : 0.0;
        const double g_ip1 = (h_ip1 != 0.0) ? 1.0 / h_ip1 : 0.0;
        off[i] = h_ip1;
        diag[i] = mul_rn(2.0, add_rn(h_ip1, h_i));
        g[i] = mul_rn(3.0, sub_rn(mul_rn(yd_ip1, g_ip1), mul_rn(yd_i, g_i)));
      }
      if (sys == 1) {
        c[1] = g[0] / diag[0];
      } else if (sys > 1) {
        double* alpha = sm.ks[4];
        double* gamma = sm.ks[5];
        double* z = diag;  // diag[i] is dead once alpha[i] is formed
        alpha[0] = diag[0];
        gamma[0] = off[0] / alpha[0];
        for (int i = 1; i < sys - 1; ++i) {
          alpha[i] = sub_rn(diag[i], mul_rn(off[i - 1], gamma[i - 1]));
          gamma[i] = off[i] / alpha[i];
        }
        alpha[sys - 1] =
            sub_rn(diag[sys - 1], mul_rn(off[sys - 2], gamma[sys - 2]));
        z[0] = g[0];
        for (int i = 1; i < sys; ++i)
          z[i] = sub_rn(g[i], mul_rn(gamma[i - 1], z[i - 1]));
        for (int i = 0; i < sys; ++i) z[i] = z[i] / alpha[i];
        c[sys] = z[sys - 1];
        for (int i = sys - 2; i >= 0; --i)
          c[i + 1] = sub_rn(z[i], mul_rn(gamma[i], c[i + 2]));
      }
      // coeff_calc per interval
      for (int i = 0; i < nk - 1; ++i) {
        const double dx = sub_rn(xa[i + 1], xa[i]);
        const double dy = sub_rn(lkv[i + 1], lkv[i]);
        sm.kn[0][i] = lkv[i];
        sm.kn[1][i] = sub_rn(dy / dx,
                             mul_rn(dx, add_rn(c[i + 1], mul_rn(2.0, c[i]))) / 3.0);
        sm.kn[2][i] = c[i];
        sm.kn[3][i] = sub_rn(c[i + 1], c[i]) / mul_rn(3.0, dx);
      }
    }
  }
  for (int i = tid; i < kM1; i += kThreads) {
    sm.w256[i] = bs.W256[i];
    sm.wlo[i] = bs.Wlo[i];
  }
  for (int i = tid; i < M2; i += kThreads) sm.wm2[i] = bs.Wm2[i];
  __syncthreads();
  if (sm.isc[I_STATUS] != POFD_OK) {
    if (tid == 0) {
      status[walker] = sm.isc[I_STATUS];
      double prev = accumulate ? loglike[walker] : 0.0;
      loglike[walker] = (sm.isc[I_STATUS] == POFD_E_BADLIKE)
                            ? prev + 1e25
                            : __longlong_as_double(0x7ff8000000000000ULL);
    }
    return;
  }
  const double sigmult = sm.sc[S_SIGMULT];

  // ---- R at the interpolation nodes (initRInterp + array getR) -----------
  double* yn = reinterpret_cast<double*>(sm.buf);  // ninterp (<= 2048)
  double* gR = yn + 2048;                          // ninterp - 2
  double* zz = gR + 2048;                          // ninterp - 2
  double* xs = zz + 2048;                          // ninterp - 2
  {
#pragma clang fp contract(off)
    for (int i = tid; i < ninterp; i += kThreads) {
      double workval = 0.0;
      for (int j = 0; j < bs.nbeam; ++j) {
        const size_t e = static_cast<size_t>(j) * ninterp + i;
        const int k = bs.rk[e];
        if (k < 0) continue;
        const double dx = bs.rdelx[e];
        const double s =
            sm.kn[0][k] + dx * (sm.kn[1][k] + dx * (sm.kn[2][k] + dx * sm.kn[3][k]));
        const double cts = exp2(s);
        if (cts > 0) workval += bs.bw[j] * cts * bs.bieta[j];
      }
      const double R = bs.prefac * workval;
      yn[i] = log2_or_small(R);
      ynode_g[i] = yn[i];
    }
  }
  __syncthreads();
  if (tid == 0) sm.tstamp[1] = wall_clock64();

  // ---- natural spline over the nodes: the tridiagonal matrix is fixed
  // (theta-independent LDL^T factors from the host); forward and backward
  // substitution as truncated recurrences.  The multipliers are < 1/2, so
  // starting kWarm = 64 steps early from zero changes no bit that matters.
  const int sys = ninterp - 2;
  {
#pragma clang fp contract(off)
    for (int i = tid; i < sys; i += kThreads)
      gR[i] = 3.0 * ((yn[i + 2] - yn[i + 1]) * bs.ginv[i + 1] -
                     (yn[i + 1] - yn[i]) * bs.ginv[i]);
    __syncthreads();
    const int per = (sys + kThreads - 1) / kThreads;
    const int lo = tid * per, hi = min(lo + per, sys);
    if (lo < hi) {
      int i0 = lo - kWarm;
      double z;
      if (i0 <= 0) {
        i0 = 0;
        z = gR[0];
      } else {
        z = gR[i0];
      }
      if (i0 >= lo) zz[i0] = z;
      for (int i = i0 + 1; i < hi; ++i) {
        z = gR[i] - bs.lgamma[i - 1] * z;
        if (i >= lo) zz[i] = z;
      }
    }
    __syncthreads();
    for (int i = lo; i < hi; ++i) zz[i] = zz[i] / bs.lalpha[i];
    __syncthreads();
    if (lo < hi) {
      int i1 = hi - 1 + kWarm;
      if (i1 > sys - 1) i1 = sys - 1;
      double x = zz[i1];
      if (i1 < hi) xs[i1] = x;
      for (int i = i1 - 1; i >= lo; --i) {
        x = zz[i] - bs.lgamma[i] * x;
        if (i < hi) xs[i] = x;
      }
    }
    __syncthreads();
    // per-interval coefficients (GSL coeff_calc), node c = xs shifted
    for (int k = tid; k < ninterp - 1; k += kThreads) {
      const double c_i = (k == 0) ? 0.0 : xs[k - 1];
      const double c_ip1 = (k + 1 == ninterp - 1) ? 0.0 : xs[k];
      const double dx = bs.xnode[k + 1] - bs.xnode[k];
      const double dy = yn[k + 1] - yn[k];
      double* cf = coef + 4 * k;
      cf[0] = yn[k];
      cf[1] = (dy / dx) - dx * (c_ip1 + 2.0 * c_i) / 3.0;
      cf[2] = c_i;
      cf[3] = (c_ip1 - c_i) / (3.0 * dx);
    }
  }

  // ---- spectral rows needed by the widest field of this beam set ---------
  if (tid == 0) {
    sm.tstamp[2] = wall_clock64();
    int kmaxb = 0;
    for (int f = 0; f < bs.nfields; ++f) {
      const int kf = spectral_kmax(sigmult * bs.fields[f].sigma_base, bs.iflux,
                                   bs.spectral_cut, M);
      kmaxb = max(kmaxb, kf);
    }
    sm.isc[I_KMAXB] = kmaxb;
    for (int r = 0; r < M2; ++r) sm.rowof[r] = -1;
    int nrows = 0;
    const int klo = min(kmaxb, M - 1);
    const int lo_rows = klo / kM1 + 1;
    int hi_start = (kmaxb >= M) ? 0 : (M - kmaxb) / kM1;
    if (kmaxb == 0) hi_start = M2;  // only Z[0]
    for (int r = 0; r < M2; ++r)
      if (r < lo_rows || r >= hi_start) {
        sm.rowof[r] = nrows;
        sm.rowk2[nrows++] = r;
      }
    sm.isc[I_NROWS] = nrows;
  }
  __syncthreads();
  const int nrows = sm.isc[I_NROWS];
  const int kmaxb = sm.isc[I_KMAXB];

  // ---- forward transform with the R fill fused in -------------------------
  double mn_part = 0.0, var_part = 0.0;
  const int nchunks = M2 / kChunk;
  const int row4 = tid >> 2, q = tid & 3;
  for (int p0 = 0; p0 < nrows; p0 += kRowsPerPass) {
    const int np = min(kRowsPerPass, nrows - p0);
    double2 acc[kRowsPerPass];
#pragma unroll
    for (int r = 0; r < kRowsPerPass; ++r) acc[r] = make_double2(0.0, 0.0);
    for (int c = 0; c < nchunks; ++c) {
      // fill: thread (n1 = row4, q) makes R at 8 consecutive grid points
      {
        const int j0 = 2 * M2 * row4 + 2 * kChunk * c + 8 * q;
        int k = bs.kst[j0 >> 3];
        double r8[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int j = j0 + e;
          double rv = 0.0;
          if (j >= bs.minit && j <= bs.maxit) {
            const double cflux = static_cast<double>(j) * dflux;
            while (k < ninterp - 2 && cflux >= bs.xnode[k + 1]) ++k;
            const double delx = cflux - bs.xnode[k];
            const double* cf = coef + 4 * k;
            const double s = cf[0] + delx * (cf[1] + delx * (cf[2] + delx * cf[3]));
            rv = exp2(s) * dflux;
          }
          r8[e] = rv;
          if (p0 == 0) {
            // getMeanVarFromR (positive R only: wrapRidx = N-1)
            if (j >= 1 && j <= bs.wrapRidx) {
              const double cf = static_cast<double>(j) * dflux;
              const double prod = (j == bs.wrapRidx) ? 0.5 * cf * rv : cf * rv;
              mn_part += prod;
              var_part += cf * prod;
            }
          }
        }
#pragma unroll
        for (int cc = 0; cc < 4; ++cc)
          sm.buf[(4 * q + cc) * kPad + row4] =
              make_double2(r8[2 * cc], r8[2 * cc + 1]);
      }
      __syncthreads();
      fft256_cols<-1>(sm.buf, sm.w256);
      // twiddle and accumulate the kept rows: thread (k1 = row4, q)
      {
        const int k1 = row4;
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) {
          const int col = 4 * q + cc;
          const int n2 = kChunk * c + col;
          double2 a = sm.buf[col * kPad + k1];
          const int t = k1 * n2;
          a = cmul(a, cmul(sm.wlo[t & (kM1 - 1)], sm.wm2[t >> 8]));
#pragma unroll
          for (int r = 0; r < kRowsPerPass; ++r)
            if (r < np) {
              const int k2 = sm.rowk2[p0 + r];
              const double2 w = sm.wm2[(n2 * k2) & (M2 - 1)];
              acc[r] = cadd(acc[r], cmul(a, w));
            }
        }
      }
      __syncthreads();
    }
    // combine the 4 partial accumulators of each k1 (lanes q = 0..3)
#pragma unroll
    for (int r = 0; r < kRowsPerPass; ++r) {
      double2 v = acc[r];
      v.x += __shfl_xor(v.x, 1);
      v.y += __shfl_xor(v.y, 1);
      v.x += __shfl_xor(v.x, 2);
      v.y += __shfl_xor(v.y, 2);
      if (q == 0 && r < np) zrows[(p0 + r) * kM1 + row4] = v;
    }
  }
  if (tid == 0) sm.tstamp[3] = wall_clock64();
  {
    const double mn = block_sum(mn_part, sm.red);
    const double var = block_sum(var_part, sm.red);
    if (tid == 0) {
      sm.sc[S_MN] = mn;
      sm.sc[S_VAR] = var;
      const double2 z0 = zrows[0];  // row k2 = 0 is always row position 0
      const double r0 = z0.x + z0.y;
      sm.sc[S_R0] = r0;
      const double sgpos = sqrt(var);
      const bool doshift = (sgpos > dflux) && (fabs(mn) > dflux);
      sm.isc[I_DOSHIFT] = doshift;
      sm.sc[S_SHIFT] = doshift ? -mn : 0.0;
    }
  }
  __syncthreads();
  const double var_noi = sm.sc[S_VAR];
  const double r0 = sm.sc[S_R0];
  {
    // r2c post-processing + prtrans = exp(R^ - R^0) e^{i(phase - shift w)}
    const double wfac = sm.sc[S_SHIFT] * bs.iflux;
    const bool doshift = sm.isc[I_DOSHIFT];
    for (int k = tid; k <= kmaxb; k += kThreads) {
      if (k == 0) {
        prtrans[0] = make_double2(1.0, 0.0);
        continue;
      }
      const int kk = (k == M) ? 0 : k;
      const int mk = (M - k) % M;
      const double2 zk = zrows[sm.rowof[kk >> 8] * kM1 + (kk & (kM1 - 1))];
      const double2 zm = zrows[sm.rowof[mk >> 8] * kM1 + (mk & (kM1 - 1))];
      const double2 A = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
      const double2 D = make_double2(zk.x - zm.x, zk.y + zm.y);
      const double2 B = make_double2(0.5 * D.y, -0.5 * D.x);
      double sn, cs;
      sincospi(static_cast<double>(k) / static_cast<double>(M), &sn, &cs);
      const double2 X = cadd(A, cmul(make_double2(cs, -sn), B));
      double expfac, ival;
      {
#pragma clang fp contract(off)
        expfac = exp(X.x - r0);
        ival = doshift ? X.y - wfac * static_cast<double>(k) : X.y;
      }
      double s2, c2;
      sincos(ival, &s2, &c2);
      prtrans[k] = make_double2(expfac * c2, expfac * s2);
    }
  }
  __syncthreads();

  // ------------------------------------------------------- datasets (fields)
  if (tid == 0) sm.tstamp[4] = wall_clock64();
  double ll_total = 0.0;
  int st_out = POFD_OK;
  for (int f = 0; f < bs.nfields; ++f) {
    const FieldDev& fd = bs.fields[f];
    double* pofd = wk.pofd + (static_cast<size_t>(walker) * bs.nfields + f) * wk.pofd_stride;
    double* bmin = wk.bmin + (static_cast<size_t>(walker) * bs.nfields + f) * wk.blk_stride;
    double* bsum = wk.bsum + (static_cast<size_t>(walker) * bs.nfields + f) * wk.blk_stride;
    double* dbg = wk.dbg + (static_cast<size_t>(walker) * bs.nfields + f) * kDbg;
    const double sigma = sigmult * fd.sigma_base;
    if (tid == 0) {
      for (int i = S_PEAK; i <= S_SHIGH; ++i) sm.sc[i] = 0.0;
      sm.isc[I_SPLIT] = -1;
      sm.isc[I_KMAXF] = spectral_kmax(sigma, bs.iflux, bs.spectral_cut, M);
      // getPD: sgpos / sgneg (src/PDFactory.cc:1225-1232)
      const double sgpos = sqrt(var_noi + sigma * sigma);
      const double sgneg = (9 * sigma * sigma > var_noi) ? sigma : sgpos;
      sm.sc[S_SGPOS] = sgpos;
      sm.sc[S_SGNEG] = sgneg;
      // findSplitPoint search range (src/PDFactory.cc:790-801)
      const double n1 = 2.5;
      const double n1sigpos = n1 * sgpos, n1signeg = n1 * sgneg;
      double bd = n1sigpos / dflux;
      bd = (bd > 2e9) ? 2e9 : bd;
      int botidx = static_cast<int>(bd);
      int topidx;
      if (n1sigpos > fabs(bs.minflux_R)) {
        topidx = bs.wrapRidx;
      } else {
        double td = -(n1signeg + bs.minflux_R) / dflux + bs.wrapRidx + 1;
        td = (td > 2e9) ? 2e9 : ((td < -2e9) ? -2e9 : td);
        topidx = static_cast<int>(td);
      }
      sm.isc[I_BOT] = botidx;
      sm.isc[I_TOP] = topidx;
    }
    __syncthreads();
    const int kmaxf = sm.isc[I_KMAXF];
    // noise factor (src/PDFactory.cc:1172-1188)
    {
      const double sigfac = -0.5 * sigma * sigma;
      for (int k = tid; k <= kmaxf; k += kThreads) {
#pragma clang fp contract(off)
        if (k == 0 || !(sigma > 0)) {
          xp[k] = (k == 0) ? make_double2(1.0, 0.0) : prtrans[k];
          continue;
        }
        const double w = bs.iflux * static_cast<double>(k);
        const double ef = exp(sigfac * w * w);
        const double2 pt = prtrans[k];
        xp[k] = make_double2(ef * pt.x, ef * pt.y);
      }
    }
    __syncthreads();
    // c2r pre-processing into the kept rows: Zp = E + i O with
    // E = X[k] + conj(X[M-k]), O = (X[k] - conj(X[M-k])) e^{+i pi k/M};
    // imaginary parts of X[0] and X[M] are ignored as FFTW's c2r does.
    double2* zp = zrows + nrows * kM1;
    for (int e = tid; e < nrows * kM1; e += kThreads) {
      const int r = e >> 8, k1 = e & (kM1 - 1);
      const int k = k1 + kM1 * sm.rowk2[r];
      double2 xa = (k <= kmaxf) ? xp[k] : make_double2(0.0, 0.0);
      const int mk = M - k;
      double2 xb = (mk <= kmaxf) ? xp[mk] : make_double2(0.0, 0.0);
      if (k == 0) {
        xa.y = 0.0;
        xb.y = 0.0;
      }
      const double2 E = make_double2(xa.x + xb.x, xa.y - xb.y);
      const double2 D = make_double2(xa.x - xb.x, xa.y + xb.y);
      double sn, cs;
      sincospi(static_cast<double>(k) / static_cast<double>(M), &sn, &cs);
      const double2 O = cmul(D, make_double2(cs, sn));
      zp[e] = make_double2(E.x - O.y, E.y + O.x);
    }
    __syncthreads();

    // ---- inverse transform, clamp, statistics, P(D) to HBM -------------
    const int botidx = sm.isc[I_BOT], topidx = sm.isc[I_TOP];
    double pk = -1.0, mnv = 0.0, s_part = 0.0, sj_part = 0.0;
    int mni = -1;
    double2 zr[kRowsPerPass];
    if (nrows <= kRowsPerPass) {
#pragma unroll
      for (int r = 0; r < kRowsPerPass; ++r)
        zr[r] = (r < nrows) ? zp[r * kM1 + row4] : make_double2(0.0, 0.0);
    }
    for (int c = 0; c < nchunks; ++c) {
      {
        const int k1 = row4;
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) {
          const int n2 = kChunk * c + 4 * q + cc;
          double2 b = make_double2(0.0, 0.0);
          if (nrows <= kRowsPerPass) {
#pragma unroll
            for (int r = 0; r < kRowsPerPass; ++r)
              if (r < nrows) {
                const double2 w = sm.wm2[(n2 * sm.rowk2[r]) & (M2 - 1)];
                b = cadd(b, cmulc(zr[r], w));
              }
          } else {
            for (int r = 0; r < nrows; ++r) {
              const double2 w = sm.wm2[(n2 * sm.rowk2[r]) & (M2 - 1)];
              b = cadd(b, cmulc(zp[r * kM1 + k1], w));
            }
          }
          const int t = k1 * n2;
          b = cmulc(b, cmul(sm.wlo[t & (kM1 - 1)], sm.wm2[t >> 8]));
          sm.buf[(4 * q + cc) * kPad + k1] = b;
        }
      }
      __syncthreads();
      fft256_cols<+1>(sm.buf, sm.w256);
      {
        const int n1 = row4;
        const int j0 = 2 * M2 * n1 + 2 * kChunk * c + 8 * q;
        double v8[8];
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) {
          const double2 z = sm.buf[(4 * q + cc) * kPad + n1];
          v8[2 * cc] = z.x;
          v8[2 * cc + 1] = z.y;
        }
        double bm = 0.0, bsm = 0.0;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          double v = v8[e];
          if (v < 0) v = 0.0;  // unwrapAndNormalizePD positivity
          v8[e] = v;
          const int j = j0 + e;
          pk = (v > pk) ? v : pk;
          if (j >= botidx && j <= topidx) {
            if (mni < 0 || v < mnv) {
              mnv = v;
              mni = j;
            }
          }
          s_part += v;
          sj_part += static_cast<double>(j) * v;
          bm = (e == 0 || v < bm) ? v : bm;
          bsm += v;
        }
        double4* dst = reinterpret_cast<double4*>(pofd + j0);
        dst[0] = make_double4(v8[0], v8[1], v8[2], v8[3]);
        dst[1] = make_double4(v8[4], v8[5], v8[6], v8[7]);
        // 32-value block statistics over the 4 lanes q = 0..3
        double o = __shfl_xor(bm, 1);
        bm = (o < bm) ? o : bm;
        o = __shfl_xor(bm, 2);
        bm = (o < bm) ? o : bm;
        bsm += __shfl_xor(bsm, 1);
        bsm += __shfl_xor(bsm, 2);
        if (q == 0) {
          const int b = j0 / kBlk;
          bmin[b] = bm;
          bsum[b] = bsm;
        }
      }
      __syncthreads();
    }
    // make this workgroup's P(D) stores visible to its own later loads
    __threadfence_block();
    if (tid == 0) sm.tstamp[5] = wall_clock64();
    {
      const double peak = block_max(pk, sm.red);
      block_argmin(mnv, mni, sm.red);
      const double S = block_sum(s_part, sm.red);
      const double SJ = block_sum(sj_part, sm.red);
      if (tid == 0) {
        sm.sc[S_PEAK] = peak;
        sm.sc[S_SUM] = S;
        sm.sc[S_SUMJ] = SJ;
        // botidx itself is the reference's starting candidate
        sm.isc[I_MINIDX] = mni;
        sm.sc[S_MINVAL] = mnv;
      }
    }
    __syncthreads();

    // ---- findSplitPoint (src/PDFactory.cc:691-921) ---------------------
    if (tid == 0) {
      int st = POFD_OK;
      int split = -1;
      const double f1 = 5e-14, f2 = 1e-5, f3 = 5;
      const double n1 = 2.5;
      const double sgpos = sm.sc[S_SGPOS], sgneg = sm.sc[S_SGNEG];
      const double peak = sm.sc[S_PEAK];
      int minidx = sm.isc[I_MINIDX];
      double minval = sm.sc[S_MINVAL];
      double jitter = 0.0;
      if (sgpos <= 0 || sgneg <= 0) st = POFD_E_SPLIT;
      const double fluxrange = N * dflux;
      if (st == POFD_OK && fluxrange < (n1 * sgpos + n1 * sgneg)) st = POFD_E_SPLIT;
      if (st == POFD_OK && peak <= 0.0) st = POFD_E_SPLIT;
      if (st == POFD_OK && (botidx >= N || topidx >= N || botidx < 0 || topidx < 0))
        st = POFD_E_SPLIT;
      if (st == POFD_OK && minidx < 0) {  // empty range: minidx = botidx
        minidx = botidx;
        minval = pofd[botidx];
      }
      if (st == POFD_OK && minval > f2 * peak) st = POFD_E_SPLIT;
      int decided = 0;
      if (st == POFD_OK) {
        const int pre_jitter_n = 17;
        if (minidx > pre_jitter_n) {
          bool mono = true;
          const int initidx = minidx - pre_jitter_n;
          double prev = pofd[initidx];
          jitter = prev;
          for (int i = initidx + 1; i < minidx; ++i) {
            const double cv = pofd[i];
            if (cv >= prev) mono = false;
            jitter += cv;
            prev = cv;
          }
          if (mono) {
            split = minidx;
            decided = 1;
          } else {
            jitter += minval;
            jitter = 1.1 * jitter / static_cast<double>(pre_jitter_n + 1);
          }
        } else {
          jitter = f1 * peak;
          if (minval > jitter) {
            split = minidx;
            decided = 1;
          }
        }
      }
      if (st == POFD_OK && !decided) {
        if (minval <= 0.0) {
          const int nminup = 3;
          int bi = (minidx < botidx) ? botidx : minidx - nminup;
          int ti = (minidx > topidx - nminup) ? topidx - nminup : minidx + nminup;
          if (bi < 0) bi = 0;
          for (int i = bi; i < ti && i < N; ++i) {
            const double cv = pofd[i];
            if (cv > minval) minval = cv;
          }
        }
        if (minval > 0) jitter = (f3 * minval) < jitter ? (f3 * minval) : jitter;
        if (botidx >= topidx) st = POFD_E_SPLIT;
        else if (pofd[topidx] < jitter) st = POFD_E_SPLIT;
      }
      sm.isc[I_STATUS] = st;
      sm.isc[I_SPLIT] = split;
      sm.isc[I_DONE] = decided;
      sm.sc[S_JITTER] = jitter;
      sm.isc[I_MINIDX] = minidx;
      sm.sc[S_MINVAL] = minval;
    }
    __syncthreads();
    if (sm.isc[I_STATUS] == POFD_OK && !sm.isc[I_DONE]) {
      // largest i in [botidx, topidx-1] with pofd[i] < jitter; split = i+1
      const double jitter = sm.sc[S_JITTER];
      const int lo = botidx, hi = topidx - 1;
      const int bl = lo / kBlk, bh = hi / kBlk;
      int found = -1;
      // top (possibly partial) block, scanned value by value
      {
        const int a = max(lo, bh * kBlk);
        int cand = -1;
        if (tid < kBlk) {
          const int i = a + tid;
          if (i <= hi && pofd[i] < jitter) cand = i;
        }
        found = block_max_int(cand, sm.red);
      }
      // whole blocks strictly between, via the block minima
      if (found < 0 && bh - 1 >= bl + 1) {
        for (int top = bh - 1; top >= bl + 1 && found < 0; top -= kThreads) {
          const int b = top - tid;
          int cb = -1;
          if (b >= bl + 1 && bmin[b] < jitter) cb = b;
          const int bstar = block_max_int(cb, sm.red);
          if (bstar >= 0) {
            int cand = -1;
            if (tid < kBlk) {
              const int i = bstar * kBlk + tid;
              if (pofd[i] < jitter) cand = i;
            }
            found = block_max_int(cand, sm.red);
          }
        }
      }
      // bottom partial block
      if (found < 0 && bl < bh) {
        int cand = -1;
        if (tid < kBlk) {
          const int i = lo + tid;
          if (i <= min(hi, bl * kBlk + kBlk - 1) && pofd[i] < jitter) cand = i;
        }
        found = block_max_int(cand, sm.red);
      }
      if (tid == 0) {
        if (found < 0)
          sm.isc[I_STATUS] = POFD_E_SPLIT;
        else
          sm.isc[I_SPLIT] = found + 1;
      }
      __syncthreads();
    }

    // ---- rotate, normalise, mean (unwrapAndNormalizePD) -----------------
    if (tid == 0) sm.tstamp[6] = wall_clock64();
    int st = sm.isc[I_STATUS];
    if (st == POFD_OK) {
      const int split = sm.isc[I_SPLIT];
      // sum of the P(D) at indices >= split: whole blocks + the partial one
      const int bs0 = split / kBlk;
      double part = 0.0;
      for (int b = bs0 + 1 + tid; b < N / kBlk; b += kThreads) part += bsum[b];
      if (tid < kBlk) {
        const int i = bs0 * kBlk + tid;
        if (i >= split && i < N) part += pofd[i];
      }
      const double Shigh = block_sum(part, sm.red);
      if (tid == 0) {
        const double S = sm.sc[S_SUM], SJ = sm.sc[S_SUMJ];
        const double plast = pofd[(split + N - 1) % N];
        const double pfirst = pofd[split];
        // trapezoid integral of the rotated array (PD::getIntegral)
        const double tot = (S - 0.5 * pfirst - 0.5 * plast) * dflux;
        const double itot = 1.0 / tot;
        // sum_i i pd[i] over the rotated array with the last weight halved
        const double Mrot = SJ - static_cast<double>(split) * S +
                            static_cast<double>(N) * (S - Shigh);
        const double msum = (Mrot - 0.5 * static_cast<double>(N - 1) * plast) * itot;
        const double tmn = msum * (dflux * dflux);
        sm.sc[S_INTEGRAL] = tot;
        sm.sc[S_ITOT] = itot;
        sm.sc[S_SHIGH] = Shigh;
        if (isinf(tmn) || isnan(tmn)) {
          sm.isc[I_STATUS] = POFD_E_MEAN;
        } else {
          sm.sc[S_MINFLUX] = -tmn;
        }
      }
      __syncthreads();
      st = sm.isc[I_STATUS];
    }

    // ---- log likelihood of the dataset (PD::getLogLike) -----------------
    if (st == POFD_OK) {
#pragma clang fp contract(off)
      const int split = sm.isc[I_SPLIT];
      const double itot = sm.sc[S_ITOT];
      const double minflux = sm.sc[S_MINFLUX];
      double part = 0.0;
      const double maxflux = minflux + static_cast<double>(N - 1) * dflux;
      const double idflux = 1.0 / dflux;
      const double maxfluxval = maxflux - dflux;
      auto Lrot = [&](int i) {
        int jj = i + split;
        if (jj >= N) jj -= N;
        return log2_or_small(pofd[jj] * itot);
      };
      if (fd.binned) {
        for (uint32_t i = tid; i < fd.nbins; i += kThreads) {
          const uint32_t ninbin = fd.bins[i];
          if (ninbin == 0) continue;
          const double cflux = fd.bincent0 + static_cast<double>(i) * fd.bindelta;
          const double nb = static_cast<double>(ninbin);
          if (cflux < minflux) {
            part += nb * Lrot(0);
          } else if (cflux > maxfluxval) {
            part += nb * Lrot(N - 1);
          } else {
            const double delt = (cflux - minflux) * idflux;
            const int idx = static_cast<int>(delt);
            const double t = delt - static_cast<double>(idx);
            part += ((1.0 - t) * Lrot(idx) + t * Lrot(idx + 1)) * nb;
          }
        }
      } else {
        // log2 of the normalised P(D) in place (unrotated order), then the
        // per-pixel interpolation; pixels are sorted so neighbouring
        // threads read neighbouring P(D) values
        for (int j = tid; j < N; j += kThreads) pofd[j] = log2_or_small(pofd[j] * itot);
        __threadfence_block();
        __syncthreads();
        auto Lr = [&](int i) {
          int jj = i + split;
          if (jj >= N) jj -= N;
          return pofd[jj];
        };
        const double L0 = Lr(0), Ln = Lr(N - 1);
        for (uint64_t i = tid; i < fd.npix; i += kThreads) {
          const double cflux = fd.pix[i];
          if (cflux <= minflux) {
            part += L0;
          } else if (cflux >= maxfluxval) {
            part += Ln;
          } else {
            const double delt = (cflux - minflux) * idflux;
            const int idx = static_cast<int>(delt);
            const double t = delt - static_cast<double>(idx);
            part += (1.0 - t) * Lr(idx) + t * Lr(idx + 1);
          }
        }
      }
      const double ll2 = block_sum(part, sm.red);
      if (tid == 0) {
        const double curr = 0.69314718055994530942 * ll2;  // log2toe
        sm.sc[S_LL] = curr;
      }
      __syncthreads();
    }
    if (tid == 0) {
      const int stf = sm.isc[I_STATUS];
      dbg[0] = sm.isc[I_SPLIT];
      dbg[1] = sm.sc[S_PEAK];
      dbg[2] = sm.isc[I_MINIDX];
      dbg[3] = sm.sc[S_MINVAL];
      dbg[4] = sm.sc[S_JITTER];
      dbg[5] = sm.sc[S_INTEGRAL];
      dbg[6] = -sm.sc[S_MINFLUX];
      dbg[7] = sm.sc[S_SGPOS];
      dbg[8] = sm.sc[S_SGNEG];
      dbg[9] = sm.sc[S_MN];
      dbg[10] = var_noi;
      dbg[11] = r0;
      dbg[12] = kmaxf;
      dbg[13] = nrows;
      dbg[14] = (stf == POFD_OK) ? sm.sc[S_LL] : 0.0;
      dbg[15] = stf;
      dbg[16] = sm.sc[S_ITOT];
      dbg[17] = sm.sc[S_MINFLUX];
      dbg[18] = (stf == POFD_OK && !fd.binned) ? 1.0 : 0.0;
      dbg[19] = kmaxb;
      dbg[20] = sm.sc[S_SUM];
      dbg[21] = sm.sc[S_SUMJ];
      dbg[22] = sm.sc[S_SHIGH];
      dbg[23] = sigma;
      sm.tstamp[7] = wall_clock64();
      for (int i = 0; i < 7; ++i)
        dbg[24 + i] = static_cast<double>(sm.tstamp[i + 1] - sm.tstamp[i]);
      dbg[31] = static_cast<double>(sm.tstamp[7] - sm.tstamp[0]);
      if (stf == POFD_OK) {
        ll_total += (sm.sc[S_LL] - fd.like_offset) / fd.like_norm;
      } else {
        st_out = stf;
      }
    }
    __syncthreads();
    if (sm.isc[I_STATUS] != POFD_OK) break;
  }
  if (tid == 0) {
    const double prev = accumulate ? loglike[walker] : 0.0;
    if (st_out != POFD_OK) {
      status[walker] = st_out;
      loglike[walker] = __longlong_as_double(0x7ff8000000000000ULL);
    } else {
      status[walker] = POFD_OK;
      loglike[walker] = prev + ll_total;
    }
  }
}

// Writes the P(D) of (walker, field) as PDFactory::getPD leaves PD::pd_:
// rotated at the split, normalised, log2 when is_log (debug / parity).
__global__ void materialize_pd_kernel(const double* pofd, int N, int split,
                                      double itot, int is_log, int already_log,
                                      double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int j = i + split;
  if (j >= N) j -= N;
  const double v = pofd[j];
  if (already_log) {
    out[i] = v;
  } else {
    const double p = v * itot;
    out[i] = is_log ? log2_or_small(p) : p;
  }
}

int launch_loglike1d(const BeamsetDev& bs, const WorkDev& wk, int beamset_idx,
                     const float* d_theta, uint32_t nwalk, uint32_t stride,
                     double* d_loglike, int32_t* d_status, int accumulate,
                     void* stream) {
  hipLaunchKernelGGL(loglike1d_kernel, dim3(nwalk), dim3(kThreads), 0,
                     static_cast<hipStream_t>(stream), bs, wk, beamset_idx,
                     d_theta, stride, d_loglike, d_status, accumulate);
  return hipGetLastError() == hipSuccess ? POFD_OK : POFD_E_INTERNAL;
}

int launch_materialize_pd(const double* d_pofd, int N, int split, double itot,
                          int is_log, int already_log, double* d_out,
                          void* stream) {
  const int nb = (N + 255) / 256;
  hipLaunchKernelGGL(materialize_pd_kernel, dim3(nb), dim3(256), 0,
                     static_cast<hipStream_t>(stream), d_pofd, N, split, itot,
                     is_log, already_log, d_out);
  return hipGetLastError() == hipSuccess ? POFD_OK : POFD_E_INTERNAL;
}

}  // namespace pofd
