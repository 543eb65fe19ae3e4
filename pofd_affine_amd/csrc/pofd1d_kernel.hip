// 1D P(D) log-likelihood kernels for gfx950.  For each beam set a batch of
// walkers runs through these launches (1024-thread workgroups); together
// they do the work of calcLikeSingle::getLogLike (src/calcLike.cc:389-443):
//
//   prep      per walker: knot spline, R at the interpolation nodes against
//             the beam histogram (numberCountsKnotsSpline::getR,
//             src/numberCountsKnotsSpline.cc:436-529), log2, natural cubic
//             spline over the nodes (PDFactory::initRInterp,
//             src/PDFactory.cc:426-527)
//   forward   per walker: R at every flux of the transform grid
//             (PDFactory::initR, :541-620) generated inside the forward
//             transform, the moments of getMeanVarFromR (:625-679),
//             exp(R^ - R^0) with the mean shift (:1094-1132) and, per
//             dataset, the noise factor (:1172-1188) applied to the rows
//             the inverse needs
//   inverse   per (walker, dataset): inverse transform, clamp, P(D) and its
//             block statistics to HBM
//   finalize  per (walker, dataset): findSplitPoint (:691-921), rotation,
//             normalisation and mean (:934-1009), the binned log likelihood
//             (src/PD.cc:557-619)
//   pixels    unbinned log likelihood (src/PD.cc:628-677), kPixBlocks
//             workgroups of 256 per (walker, dataset)
//   combine   (curr - like_offset)/like_norm summed over the datasets in
//             order (src/calcLike.cc:425), into the batch outputs
//
// Transform design.  The real transform of length N = 2M is a complex
// transform of length M = 256 * M2 in four-step form (m = M2*n1 + n2,
// k = k1 + 256*k2): 256-point column transforms in LDS, a twiddle, then the
// transform across columns.  After the noise factor only k <= kmax survive
// (exp(-sigma^2 w^2 / 2) below the configured cut), so the across-column
// transform is evaluated only for the rows k2 holding k <= kmax or
// k >= M - kmax, as a direct sum accumulated while the columns stream
// through LDS; the inverse mirrors it.  R values are generated in the
// forward column loop and never stored, so the only HBM traffic of size N
// per walker and dataset is the P(D) itself.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "internal.h"

namespace pofd {

namespace {

constexpr int kWaves = kThreads / 64;
// LDS layout of a chunk: element e of column col (16 columns of 256 double2)
// sits at cpos(col, e) = col * kSlots + (slot(e) ^ (col >> 2)), with
// slot(e) = e ^ ((e >> 1) & 15) ^ ((e >> 2) & 15) a permutation inside each
// block of 16 and kSlots = 257.  The layout was chosen against per-dispatch
// LDS bank-conflict counts of these access patterns on gfx950, which fit
// ds_write_b128 serving 8 consecutive lanes a pass over 8 16-byte bank
// groups and ds_read_b128 the lane sets {0-3, 20-23}, {4-7, 16-19},
// {8-11, 28-31}, {12-15, 24-27} (and the same +32) over 16 groups; under
// that model every stage read
// (e = j + 64m) and stage write (e = 4j + m, 16(j/4) + j%4 + 4m, ...) of the
// column transforms and the fill's and accumulation's column accesses
// (2 rows x 4 columns 4q + cc per pass) touch distinct groups.
constexpr int kSlots = kM1 + 1;
__device__ __forceinline__ int slot(int e) { return e ^ ((e >> 1) & 15) ^ ((e >> 2) & 15); }
__device__ __forceinline__ int cpos(int col, int e) {
  return col * kSlots + (slot(e) ^ (col >> 2));
}

// accumulator of v_mfma_f64_16x16x4_f64 (four doubles per lane)
typedef double f64x4 __attribute__((ext_vector_type(4)));

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

// The reference's scalar arithmetic is reproduced without contraction so
// that a*b+c rounds as the reference's (non-FMA) build does.
__device__ __forceinline__ double mul_rn(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ double add_rn(double a, double b) { return __dadd_rn(a, b); }
__device__ __forceinline__ double sub_rn(double a, double b) { return __dsub_rn(a, b); }

__device__ __forceinline__ double log2_or_small(double v) {
  return (v > 0.0) ? log2(v) : -30.0;  // pofd_mcmc::smalllogval
}

// lane `lane`'s v (lane uniform across the wavefront)
__device__ __forceinline__ double readlane_d(const double v, const int lane) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
  const unsigned lo = static_cast<unsigned>(
      __builtin_amdgcn_readlane(static_cast<int>(static_cast<unsigned>(b)), lane));
  const unsigned hi = static_cast<unsigned>(
      __builtin_amdgcn_readlane(static_cast<int>(static_cast<unsigned>(b >> 32)), lane));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

__device__ __forceinline__ double qnan() {
  return __longlong_as_double(0x7ff8000000000000ULL);
}

// per-walker scalars handed between the launches (WorkDev::state)
enum {
  ST_STATUS = 0, ST_SIGMULT, ST_MN, ST_VAR, ST_SHIFT, ST_DOSHIFT, ST_KMAXB,
  ST_NROWS, ST_R0, ST_TPREP, ST_TFWD, ST_TSPEC
};

// ---- block-wide reductions (deterministic fixed trees) -------------------
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
  return s;
}

// two block_sum reductions of a kThreads workgroup in one pass (red holds
// 2 kWaves values), each in block_sum's order
__device__ void block_sum2(double& a, double& b, double* red) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o);
    b += __shfl_xor(b, o);
  }
  __syncthreads();
  if ((tid & 63) == 0) {
    red[tid >> 6] = a;
    red[kWaves + (tid >> 6)] = b;
  }
  __syncthreads();
  a = red[0];
  b = red[kWaves];
#pragma unroll
  for (int i = 1; i < kWaves; ++i) {
    a += red[i];
    b += red[kWaves + i];
  }
}

// first chunk of chunk group g (of kGroups) over nchunks chunks
__device__ __forceinline__ int group_chunk(const int g, const int nchunks) {
  return g * nchunks / kGroups;
}

// Thread 0's step of the group tree (G0 + G1) + (G2 + G3) of two block
// sums a and b, taken as the groups complete in order: the open pair
// becomes G_g (g even) or G_(g-1) + G_g (g odd), and a workgroup holding
// all four groups keeps G0 + G1 when g = 1.  The open pair starts at zero
// (tree_start), so a part holding only G1 or G3 gives that group's sum.
__device__ __forceinline__ void tree_start(double* gt) {
  if (threadIdx.x == 0) gt[0] = gt[1] = gt[2] = gt[3] = 0.0;
}
__device__ __forceinline__ void tree_add(double* gt, const double a, const double b,
                                         const int g, const int parts) {
  if (threadIdx.x != 0) return;
  gt[0] = (g & 1) ? gt[0] + a : a;
  gt[1] = (g & 1) ? gt[1] + b : b;
  if (parts == 1 && g == 1) {
    gt[2] = gt[0];
    gt[3] = gt[1];
  }
}

__device__ double block_max(double v, double* red) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double u = __shfl_xor(v, o);
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
    const double u = __shfl_xor(v, o);
    const int ui = __shfl_xor(idx, o);
    const bool take = (ui >= 0) && (idx < 0 || u < v || (u == v && ui < idx));
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
    const double u = red[i];
    const int ui = static_cast<int>(red[kWaves + i]);
    const bool take = (ui >= 0) && (bi < 0 || u < bv || (u == bv && ui < bi));
    bv = take ? u : bv;
    bi = take ? ui : bi;
  }
  v = bv;
  idx = bi;
}

// The 32-value block statistics of the P(D): block b = j / kBlk (kBlk = 2
// kChunk, so row n1 = b / nbc and chunk c = b % nbc of the inverse transform,
// nbc = M2 / kChunk) is stored at c * kM1 + n1, so that the 16 rows of a
// wavefront's epilogue store whole 128-byte lines
static_assert(kBlk == 2 * kChunk, "a block is one row of one chunk");
__device__ __forceinline__ int blk_slot(const int b, const int nbc) {
  return (b % nbc) * kM1 + b / nbc;
}

// ---- 256-point transforms of 16 swizzled LDS columns ----------------------
// Radix-4 Stockham between two buffers: 64 threads (one wavefront) per
// column, 4 points each; after the four stages the result is in `a`.
// The caller synchronises before the call.  SIGN = -1 is FFTW_FORWARD, +1
// the unnormalised backward direction.  The twiddles of stage Ns come from
// wst (kStageTw entries per multiplier m = 1..3): wst[(m - 1) * kStageTw +
// stw_base(Ns) + jm] = exp(-2 pi i m jm (64 / Ns) / 256), so the lanes of a
// stage read consecutive entries.
constexpr int kStageTw = 4 + 16 + 64;
__device__ __forceinline__ int stw_base(int Ns) { return Ns == 4 ? 0 : (Ns == 16 ? 4 : 20); }

// The 20 LDS element offsets of one thread's column transforms (the four
// reads e = j + 64 m, the same in every stage, then four writes per stage),
// two 16-bit offsets per word (they are below 16 x 257): held across the
// forward's chunk loop in 10 registers instead of 20.
struct ColOffsets {
  uint32_t w[10];
  __device__ __forceinline__ int get(const int k) const {
    return static_cast<int>((w[k >> 1] >> (16 * (k & 1))) & 0xffffu);
  }
};

__device__ ColOffsets col_offsets() {
  const int tid = threadIdx.x, col = tid >> 6, j = tid & 63;
  int o[20];
#pragma unroll
  for (int m = 0; m < 4; ++m) o[m] = cpos(col, j + 64 * m);
  int k = 4;
#pragma unroll
  for (int Ns = 1; Ns < kM1; Ns *= 4) {
    const int d = (j / Ns) * Ns * 4 + (j & (Ns - 1));
#pragma unroll
    for (int q = 0; q < 4; ++q) o[k++] = cpos(col, d + q * Ns);
  }
  ColOffsets c;
#pragma unroll
  for (int i = 0; i < 10; ++i)
    c.w[i] = static_cast<uint32_t>(o[2 * i]) | (static_cast<uint32_t>(o[2 * i + 1]) << 16);
  return c;
}

// co: the packed offsets (forward), or null to compute them here (inverse)
template <int SIGN>
__device__ void fft256_cols(double2* a, double2* b, const double2* wst,
                            const ColOffsets* co = nullptr) {
  const int tid = threadIdx.x;
  const int col = tid >> 6, j = tid & 63;
  double2* src = a;
  double2* dst = b;
  int stage = 0;
#pragma unroll
  for (int Ns = 1; Ns < kM1; Ns *= 4, ++stage) {
    double2 v0, v1, v2, v3;
    if (co) {
      v0 = src[co->get(0)];
      v1 = src[co->get(1)];
      v2 = src[co->get(2)];
      v3 = src[co->get(3)];
    } else {
      v0 = src[cpos(col, j)];
      v1 = src[cpos(col, j + 64)];
      v2 = src[cpos(col, j + 128)];
      v3 = src[cpos(col, j + 192)];
    }
  const int jm = j & (Ns - 1);
    if (Ns > 1) {
      const int e = stw_base(Ns) + jm;
      double2 t1 = wst[e], t2 = wst[kStageTw + e], t3 = wst[2 * kStageTw + e];
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
    double2 y1, y3;
    if (SIGN < 0) {  // a1 -/+ i a3
      y1 = make_double2(a1.x + a3.y, a1.y - a3.x);
      y3 = make_double2(a1.x - a3.y, a1.y + a3.x);
    } else {
      y1 = make_double2(a1.x - a3.y, a1.y + a3.x);
      y3 = make_double2(a1.x + a3.y, a1.y - a3.x);
    }
    if (co) {
      dst[co->get(4 + 4 * stage)] = cadd(a0, a2);
      dst[co->get(5 + 4 * stage)] = y1;
      dst[co->get(6 + 4 * stage)] = csub(a0, a2);
      dst[co->get(7 + 4 * stage)] = y3;
    } else {
      const int d = (j / Ns) * Ns * 4 + jm;
      dst[cpos(col, d)] = cadd(a0, a2);
      dst[cpos(col, d + Ns)] = y1;
      dst[cpos(col, d + 2 * Ns)] = csub(a0, a2);
      dst[cpos(col, d + 3 * Ns)] = y3;
    }
    // a column belongs to one wavefront (64 threads) in every stage, so
    // between stages only that wavefront's LDS accesses need ordering; the
    // last stage ends with a workgroup barrier for the callers, which read
    // the columns across wavefronts
    if (Ns < kM1 / 4) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
      __syncthreads();
    }
    double2* t = src;
    src = dst;
    dst = t;
  }
}

// Largest spectral index whose noise factor survives the cut:
// sigfac*w*w >= -cut with w = iflux*k (src/PDFactory.cc:1180-1182).
__device__ int spectral_kmax(double sigma, double iflux, double cut, int M) {
  if (!(sigma > 0)) return M;
  const double sigfac = -0.5 * sigma * sigma;
  const double est = sqrt(2.0 * cut) / (sigma * iflux);
  long k = (est > static_cast<double>(M)) ? M : static_cast<long>(est);
  auto keep = [&](long kk) {
    const double w = mul_rn(iflux, static_cast<double>(kk));
    return mul_rn(mul_rn(sigfac, w), w) >= -cut;
  };
  while (k < M && keep(k + 1)) ++k;
  while (k > 0 && !keep(k)) --k;
  return static_cast<int>(k);
}

// Spectral rows k2 needed by the widest dataset of the beam set (k <= kmaxb
// or k >= M - kmaxb); called by every thread of the workgroup.
__device__ int spectral_rows(const BeamsetDev& bs, double sigmult, short* rowk2,
                             short* rowof, int& kmaxb_out) {
  const int M = bs.M, M2 = bs.M2;
  int kmaxb = 0;
  for (int f = 0; f < bs.nfields; ++f)
    kmaxb = max(kmaxb, spectral_kmax(sigmult * bs.fields[f].sigma_base,
                                     bs.iflux, bs.spectral_cut, M));
  const int klo = min(kmaxb, M - 1);
  const int lo_rows = klo / kM1 + 1;
  int hi_start = (kmaxb >= M) ? 0 : (M - kmaxb) / kM1;
  if (kmaxb == 0) hi_start = M2;  // only Z[0]
  // the kept rows are [0, lo_rows) then [h0, M2), in ascending order; every
  // thread of the workgroup computes the same bounds and fills its share
  // of the tables (the caller synchronises before reading them)
  const int h0 = max(hi_start, lo_rows);
  for (int r = threadIdx.x; r < M2; r += blockDim.x) {
    const int pos = (r < lo_rows) ? r : ((r >= h0) ? lo_rows + (r - h0) : -1);
    rowof[r] = static_cast<short>(pos);
    if (pos >= 0) rowk2[pos] = static_cast<short>(r);
  }
  kmaxb_out = kmaxb;
  return lo_rows + (M2 - h0);
}

__device__ __forceinline__ int batch_n(const WorkDev& wk) {
  return wk.dcount ? *wk.dcount : wk.nhost;
}

__device__ __forceinline__ double* walker_state(const WorkDev& wk, int w) {
  return wk.state + static_cast<size_t>(w) * kState;
}

// ---- LDS layouts -----------------------------------------------------------
constexpr int kBeamStage = 1024;
struct PrepSmem {
  double yn[2048];   // log2 R at the nodes
  double gR[2048];   // right-hand side of the node system
  double zz[2048];   // forward substitution, then scaled by 1/alpha
  double xs[2048];   // solution (node c values shifted by one)
  double lg[2048];   // LDL^T multipliers
  double la[2048];   // LDL^T diagonal
  double gi[2048];   // 1/h
  double kn[4][kMaxKnots];
  double bw[kBeamStage];  // beam weights and inverse values (staged when
  double be[kBeamStage];  // nbeam + nbeamn <= kBeamStage)
  double red[kWaves];
  int isc[4];
};

struct XformSmem {
  double2 bufa[kChunk * kSlots];
  double2 bufb[kChunk * kSlots];
  double2 wst[3 * kStageTw];  // stage twiddles of the column transforms
  double2 wlo[kM1];
  double2 wm2[512];
  short rowk2[512];
  short rowof[512];
  double red[kWaves * 2];
  double sc[4];
  // the group tree of two block sums (kept by thread 0): the open pair of
  // each, then the closed G0 + G1 of each
  double gtree[4];
  double ex2[32];           // 2^(i/32) * dflux (forward only)
  int isc[2 + kMaxFields];  // nrows, kmaxb, then kmax of each dataset
};

__device__ void load_twiddles(const BeamsetDev& bs, XformSmem& sm) {
  for (int i = threadIdx.x; i < kM1; i += blockDim.x) sm.wlo[i] = bs.Wlo[i];
  for (int i = threadIdx.x; i < 3 * kStageTw; i += blockDim.x) {
    const int m = i / kStageTw + 1, r = i % kStageTw;
    const int Ns = (r < 4) ? 4 : (r < 20 ? 16 : 64);
    const int jm = r - stw_base(Ns);
    sm.wst[i] = bs.W256[(m * jm * (64 / Ns)) & (kM1 - 1)];
  }
  for (int i = threadIdx.x; i < bs.M2; i += blockDim.x) sm.wm2[i] = bs.Wm2[i];
}

// 2^s * scale by Tang's table-driven method (ACM TOMS 15:144, 1989):
// 32 s = 32 n + j + 32 r with |r| <= 1/64, 2^s = 2^n 2^(j/32) 2^r, the table
// entries tab[j] = 2^(j/32) * scale (the fill's scale is dflux) and
// 2^r = exp(r ln 2) as its Taylor polynomial of degree 6 (truncation below
// 4e-18 relative).  Accurate to about one ulp; 0 below 2^-1075 and +inf
// above 2^1024 as exp2 gives them.
__device__ __forceinline__ double exp2_dflux(double s, const double* tab) {
  const double t = fmin(fmax(s * 32.0, -35000.0), 33000.0);
  const double n = rint(t);
  const double r = (t - n) * 0.03125;
  const int ni = static_cast<int>(n);
  double p = 1.5403530393381609954e-4;   // (ln 2)^6 / 720
  p = fma(p, r, 1.3333558146428443423e-3);  // (ln 2)^5 / 120
  p = fma(p, r, 9.6181291076284771620e-3);  // (ln 2)^4 / 24
  p = fma(p, r, 5.5504108664821579953e-2);  // (ln 2)^3 / 6
  p = fma(p, r, 2.4022650695910071233e-1);  // (ln 2)^2 / 2
  p = fma(p, r, 6.9314718055994530942e-1);  // ln 2
  p = fma(p, r, 1.0);
  return ldexp(tab[ni & 31] * p, ni >> 5);
}

}  // namespace

// ===========================================================================
// prep: knot spline, R at the nodes, node spline, interval coefficients
// ===========================================================================
// The natural spline over the nodes of segment seg from log2 R in sm.yn and
// the segment's theta-independent factors in sm.lg / sm.la / sm.gi, into the
// walker's per-interval coefficients.  The tridiagonal matrix is fixed
// (LDL^T factors from the host); forward and backward substitution run as
// truncated recurrences, each thread's started bs.kwarm[seg] steps early:
// what the start leaves out is scaled by the product of that many
// multipliers, at most 2^-64 (the multipliers are < 1/2; at most 64 steps).
// Called by every thread.
__device__ void node_spline(const BeamsetDev& bs, const WorkDev& wk, const int walker,
                            const int seg, PrepSmem& sm) {
#pragma clang fp contract(off)
  const int tid = threadIdx.x;
  const int ninterp = bs.ninterp, sys = ninterp - 2;
  const double* xnode = bs.xnode + seg * ninterp;
  for (int i = tid; i < sys; i += kThreads)
    sm.gR[i] = 3.0 * ((sm.yn[i + 2] - sm.yn[i + 1]) * sm.gi[i + 1] -
                      (sm.yn[i + 1] - sm.yn[i]) * sm.gi[i]);
  __syncthreads();
  const int per = (sys + kThreads - 1) / kThreads;
  const int lo = tid * per, hi = min(lo + per, sys);
  const int kw = bs.kwarm[seg];
  // The recurrences' operands are read from LDS four steps ahead, so that
  // a step waits only for the one before it (the steps and their order are
  // those of the plain loops).
  if (lo < hi) {
    int i0 = lo - kw;
    double z;
    if (i0 <= 0) {
      i0 = 0;
      z = sm.gR[0];
    } else {
      z = sm.gR[i0];
    }
    if (i0 >= lo) sm.zz[i0] = z;
    int i = i0 + 1;
    for (; i + 4 <= lo; i += 4) {  // warm-up: no stores
      double g4[4], l4[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        g4[u] = sm.gR[i + u];
        l4[u] = sm.lg[i + u - 1];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) z = g4[u] - l4[u] * z;
    }
    for (; i < lo; ++i) z = sm.gR[i] - sm.lg[i - 1] * z;
    for (; i < hi; ++i) {
      z = sm.gR[i] - sm.lg[i - 1] * z;
      sm.zz[i] = z;
    }
  }
  __syncthreads();
  for (int i = lo; i < hi; ++i) sm.zz[i] = sm.zz[i] / sm.la[i];
  __syncthreads();
  if (lo < hi) {
    int i1 = hi - 1 + kw;
    if (i1 > sys - 1) i1 = sys - 1;
    double x = sm.zz[i1];
    if (i1 < hi) sm.xs[i1] = x;
    int i = i1 - 1;
    for (; i - 3 >= hi; i -= 4) {  // warm-up: no stores
      double z4[4], l4[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        z4[u] = sm.zz[i - u];
        l4[u] = sm.lg[i - u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) x = z4[u] - l4[u] * x;
    }
    for (; i >= hi; --i) x = sm.zz[i] - sm.lg[i] * x;
    for (; i >= lo; --i) {
      x = sm.zz[i] - sm.lg[i] * x;
      sm.xs[i] = x;
    }
  }
  __syncthreads();
  // per-interval coefficients (GSL coeff_calc), node c = xs shifted
  double* coef = wk.coef + walker * wk.coef_stride + 4 * seg * ninterp;
  for (int k = tid; k < ninterp - 1; k += kThreads) {
    const double c_i = (k == 0) ? 0.0 : sm.xs[k - 1];
    const double c_ip1 = (k + 1 == ninterp - 1) ? 0.0 : sm.xs[k];
    const double dx = xnode[k + 1] - xnode[k];
    const double dy = sm.yn[k + 1] - sm.yn[k];
    double* cf = coef + 4 * k;
    cf[0] = sm.yn[k];
    cf[1] = (dy / dx) - dx * (c_ip1 + 2.0 * c_i) / 3.0;
    cf[2] = c_i;
    cf[3] = (c_ip1 - c_i) / (3.0 * dx);
  }
  __syncthreads();  // LDS reused by the next segment
}

// the segment's theta-independent node-spline factors into LDS
// (by threads t0, t0 + nt, ...: the whole workgroup by default)
__device__ void load_node_factors(const BeamsetDev& bs, const int seg, PrepSmem& sm,
                                  const int t0 = threadIdx.x, const int nt = kThreads) {
  const int ninterp = bs.ninterp, sys = ninterp - 2;
  for (int i = t0; i < sys; i += nt) {
    sm.lg[i] = bs.lgamma[seg * sys + i];
    sm.la[i] = bs.lalpha[seg * sys + i];
  }
  for (int i = t0; i < ninterp - 1; i += nt) sm.gi[i] = bs.ginv[seg * (ninterp - 1) + i];
}

// A small batch spreads each walker over `parts` workgroups (the split
// transforms' count): part p takes the nodes [p n / parts, (p + 1) n /
// parts) of each segment and writes their log2 R to the workspace;
// prepspline1d_kernel then makes the node splines.  With one part the
// workgroup does both.
__global__ __launch_bounds__(kThreads) void prep1d_kernel(
    const BeamsetDev bs, const WorkDev wk, const float* __restrict__ theta,
    const uint32_t stride, const int32_t* __restrict__ prev_status,
    const int accumulate) {
  __shared__ PrepSmem sm;
  const int tid = threadIdx.x;
  const int n = batch_n(wk);
  const int parts = batch_parts(n, wk.force_parts, wk.nslots, wk.parts_tab);
  const int walker = blockIdx.x / parts;
  const int part = blockIdx.x - walker * parts;
  if (walker >= n) return;
  const int ninterp = bs.ninterp, nk = bs.nknots;
  double* st = walker_state(wk, walker);
  const unsigned long long t0 = wall_clock64();

  // the first node segment's fixed factors, loaded while wave 0 makes the
  // knot spline
  const int seg0 = bs.haspos ? 0 : 1;
  if (parts == 1 && tid >= 64) load_node_factors(bs, seg0, sm, tid - 64, kThreads - 64);
  // and the beam weights and inverse values, when they fit in LDS
  const bool staged = bs.nbeam + bs.nbeamn <= kBeamStage;
  if (staged) {
    for (int j = tid - 64; j >= 0 && j < bs.nbeam + bs.nbeamn; j += kThreads - 64) {
      sm.bw[j] = bs.bw[j];
      sm.be[j] = bs.bieta[j];
    }
  }
  if (tid < 64) {
    // The knot spline on the 64 lanes of wave 0 (nknots <= 64), lane l for
    // knot or interval l: GSL's natural cubic spline through (logknots,
    // logknotvals), cspline_init + gsl_linalg_solve_symm_tridiag +
    // coeff_calc, with the knot positions' factors (h, 1/h, LDL^T diagonal
    // and multipliers) from the host, rounded as GSL's non-FMA build rounds
    // them; the two substitutions run in order over broadcast lane values.
    const int l = tid;
    const float* th = theta + static_cast<size_t>(walker) * stride;
    // numberCountsKnotsSpline::setParams src/numberCountsKnotsSpline.cc:265-274
    const double lk = (l < nk) ? mul_rn(bs.logfac, static_cast<double>(th[l])) : 0.0;
    const bool valid = __ballot(l < nk && isnan(lk)) == 0;
    const int sys = nk - 2;
    const double* kf = bs.kfac;
    const double hI = kf[l], giI = kf[kMaxKnots + l];
    const double alI = kf[2 * kMaxKnots + l], gmI = kf[3 * kMaxKnots + l];
    const double lk1 = __shfl_down(lk, 1), lk2 = __shfl_down(lk, 2);
    const double giI1 = __shfl_down(giI, 1);
    // right-hand side g_i (lanes i < sys)
    const double ydI = sub_rn(lk1, lk), ydI1 = sub_rn(lk2, lk1);
    const double gI = mul_rn(3.0, sub_rn(mul_rn(ydI1, giI1), mul_rn(ydI, giI)));
    // z_0 = g_0, z_i = g_i - gamma_(i-1) z_(i-1); then z_i / alpha_i
    double zl = 0.0, z = 0.0;
    for (int i = 0; i < sys; ++i) {
      const double gi = readlane_d(gI, i);
      z = (i == 0) ? gi : sub_rn(gi, mul_rn(readlane_d(gmI, i - 1), z));
      if (l == i) zl = z;
    }
    if (l < sys) zl = zl / alI;
    // c_sys = z_(sys-1), c_(i+1) = z_i - gamma_i c_(i+2); c_0 = c_(nk-1) = 0
    double cl = 0.0;
    if (sys >= 1) {
      double cc = readlane_d(zl, sys - 1);
      if (l == sys) cl = cc;
      for (int i = sys - 2; i >= 0; --i) {
        cc = sub_rn(readlane_d(zl, i), mul_rn(readlane_d(gmI, i), cc));
        if (l == i + 1) cl = cc;
      }
    }
    const double cl1 = __shfl_down(cl, 1);
    if (valid && l < nk - 1) {
      const double dx = hI;
      const double dy = sub_rn(lk1, lk);
      sm.kn[0][l] = lk;
      sm.kn[1][l] = sub_rn(dy / dx, mul_rn(dx, add_rn(cl1, mul_rn(2.0, cl))) / 3.0);
      sm.kn[2][l] = cl;
      sm.kn[3][l] = sub_rn(cl1, cl) / mul_rn(3.0, dx);
    }
    if (l == 0) {
      int status = accumulate ? prev_status[walker] : POFD_OK;
      const double sigmult = static_cast<double>(th[nk]);
      if (status == POFD_OK) {
        if (isnan(sigmult))
          status = POFD_E_ARG;
        else if (mul_rn(bs.maxsigma_base, sigmult) < 0.0)
          status = POFD_E_BADLIKE;  // calcLikeSingle::bad_like, calcLike.cc:403
        else if (!valid)
          status = POFD_E_INVALID_MODEL;  // PDFactory.cc:1046-1047
      }
      if (part == 0) {
        st[ST_STATUS] = status;
        st[ST_SIGMULT] = sigmult;
      }
      sm.isc[0] = status;
      sm.isc[1] = valid ? 1 : 0;
    }
  }
  __syncthreads();
  if (bs.bonus && part == 0) {
    // getNS / getFluxPerArea / getFluxSqPerArea (numberCountsKnotsSpline::
    // splineInt, src/numberCountsKnotsSpline.cc:300-343): int S^a n(S) dS
    // over the knot range, a = 0, 1, 2, as Gauss-Legendre sums in
    // v = log2 S over each knot interval (the integrand 2^(spline + (a+1) v)
    // ln 2 is analytic there); NaN for an invalid model
    double b0 = 0.0, b1 = 0.0, b2 = 0.0;
    if (sm.isc[1]) {
      const int npts = (nk - 1) * kGL;
      for (int t = tid; t < npts; t += kThreads) {
        const int k = t / kGL, g = t - k * kGL;
        const double v0 = bs.logknots[k], v1 = bs.logknots[k + 1];
        const double h = 0.5 * (v1 - v0), c = 0.5 * (v1 + v0);
        const double v = c + h * bs.glx[g];
        const double dx = v - v0;
        const double sv =
            sm.kn[0][k] + dx * (sm.kn[1][k] + dx * (sm.kn[2][k] + dx * sm.kn[3][k]));
        const double sflux = exp2(v);
        const double f = bs.glw[g] * h * exp2(sv + v);
        b0 += f;
        b1 += f * sflux;
        b2 += f * sflux * sflux;
      }
    }
    b0 = block_sum(b0, sm.red);
    b1 = block_sum(b1, sm.red);
    b2 = block_sum(b2, sm.red);
    if (tid == 0) {
      const double ln2 = 0.69314718055994530942;
      const double nan = __longlong_as_double(0x7ff8000000000000ULL);
      double* bo = wk.bonus + static_cast<size_t>(walker) * 4;
      bo[0] = sm.isc[1] ? ln2 * b0 : nan;
      bo[1] = sm.isc[1] ? ln2 * b1 : nan;
      bo[2] = sm.isc[1] ? ln2 * b2 : nan;
    }
  }
  if (sm.isc[0] != POFD_OK) return;  // uniform across the workgroup

  // one pass per node segment: the positive beam, then the negative beam
  const int node_lo = part * ninterp / parts, node_hi = (part + 1) * ninterp / parts;
  for (int seg = seg0; seg < bs.nseg; ++seg) {
    if (parts == 1 && seg != seg0) load_node_factors(bs, seg, sm);

    // ---- R at the interpolation nodes (initRInterp + array getR) ---------
    // (the beam weights and inverse values from LDS when they fit there:
    // read from global memory inside a term they would stall the term, and
    // with it the next group's table loads, on their latency)
    {
      const int j0 = (seg == 0) ? 0 : bs.nbeam;
      const int j1 = (seg == 0) ? bs.nbeam : bs.nbeam + bs.nbeamn;
      double* ynode_g = wk.ynode + walker * wk.ynode_stride + seg * ninterp;
      auto rnodes = [&](auto staged) {
#pragma clang fp contract(off)
        constexpr bool kStaged = decltype(staged)::value;
        for (int i = node_lo + tid; i < node_hi; i += kThreads) {
          double workval = 0.0;
          // beam bins in groups of 8, the table loads of the next group
          // issued before the current group's terms are added (in bin order)
          int kq[8], kn8[8];
          double dq[8], dn8[8];
          auto load_group = [&](const int jg, int* kk, double* dd) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const int j = jg + u;
              kk[u] = -1;
              dd[u] = 0.0;
              if (j < j1) {  // uniform across the workgroup
                const size_t e = static_cast<size_t>(j) * ninterp + i;
                kk[u] = bs.rk[e];
                dd[u] = bs.rdelx[e];
              }
            }
          };
          if (j0 < j1) load_group(j0, kq, dq);
          for (int jg = j0; jg < j1; jg += 8) {
            if (jg + 8 < j1) load_group(jg + 8, kn8, dn8);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const int k = kq[u];
              if (k < 0) continue;
              const int j = jg + u;
              const double dx = dq[u];
              const double sv =
                  sm.kn[0][k] + dx * (sm.kn[1][k] + dx * (sm.kn[2][k] + dx * sm.kn[3][k]));
              const double cts = exp2(sv);
              const double wj = kStaged ? sm.bw[j] : bs.bw[j];
              const double ej = kStaged ? sm.be[j] : bs.bieta[j];
              if (cts > 0) workval += wj * cts * ej;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              kq[u] = kn8[u];
              dq[u] = dn8[u];
            }
          }
          const double R = bs.prefac * workval;
          sm.yn[i] = log2_or_small(R);
          ynode_g[i] = sm.yn[i];
        }
      };
      if (staged) {
        rnodes(std::true_type{});
      } else {
        rnodes(std::false_type{});
      }
    }
    __syncthreads();
    if (parts == 1) node_spline(bs, wk, walker, seg, sm);
  }
  if (tid == 0 && part == 0) st[ST_TPREP] = static_cast<double>(wall_clock64() - t0);
}

// the node splines of a split batch (one workgroup per walker), from the
// log2 R the parts wrote
__global__ __launch_bounds__(kThreads) void prepspline1d_kernel(const BeamsetDev bs,
                                                                 const WorkDev wk) {
  __shared__ PrepSmem sm;
  const int n = batch_n(wk);
  const int walker = blockIdx.x;
  if (walker >= n || batch_parts(n, wk.force_parts, wk.nslots, wk.parts_tab) == 1) return;
  if (walker_state(wk, walker)[ST_STATUS] != POFD_OK) return;  // uniform
  const int ninterp = bs.ninterp;
  for (int seg = bs.haspos ? 0 : 1; seg < bs.nseg; ++seg) {
    load_node_factors(bs, seg, sm);
    const double* ynode_g = wk.ynode + walker * wk.ynode_stride + seg * ninterp;
    for (int i = threadIdx.x; i < ninterp; i += kThreads) sm.yn[i] = ynode_g[i];
    __syncthreads();
    node_spline(bs, wk, walker, seg, sm);
  }
}

// ===========================================================================
// forward: R fill fused into the row-pruned forward transform, then the
// spectrum exp(R^ - R^0) and, per dataset, the noisy rows of the inverse
// ===========================================================================
__device__ void forward_tail(const BeamsetDev& bs, const WorkDev& wk, const int walker,
                             XformSmem& sm, const double mn, const double var,
                             unsigned long long tmark);

// R at grid point j, as the fill makes it (the node spline of the interval
// kst[j], exponentiated by exp2_dflux); 0 outside the fill ranges
__device__ double r_grid(const BeamsetDev& bs, const double* __restrict__ coef, const int j,
                         const double* ex2) {
  const bool inneg = (j >= bs.minitn && j <= bs.maxitn);
  if (!inneg && !(j >= bs.minit && j <= bs.maxit)) return 0.0;
  const int k = bs.kst[j];
  const double xk = bs.xnode[k];
  const double4 cf = *reinterpret_cast<const double4*>(coef + 4 * k);
  const double cflux = inneg ? static_cast<double>(j) * bs.dflux + bs.mval
                             : static_cast<double>(j) * bs.dflux;
  const double delx = cflux - xk;
  const double sv = cf.x + delx * (cf.y + delx * (cf.z + delx * cf.w));
  return exp2_dflux(sv, ex2);
}

// head sums work in blocks of this many consecutive grid points
constexpr int kHeadBlk = 8;

// E(t) = exp(-2 pi i t/N) - 1 from two small tables in LDS: t is folded to
// v in [0, N/2] (E(N - v) = conj E(v)) and split v = kEmHi h + l, and
// E(v) = E(l) + E(kEmHi h) + E(l) E(kEmHi h) exactly.  On [0, N/2] |E| grows
// with the angle, so neither part exceeds |E(v)| much and the sum has no
// cancellation: about two roundings of |E(v)|.
constexpr int kEmHi = 512;
__device__ __forceinline__ double2 em1_at(const BeamsetDev& bs, const double2* elo,
                                          const double2* ehi, uint32_t t) {
  const int u = static_cast<int>(t & static_cast<uint32_t>(bs.N - 1));
  const bool lo = u <= bs.M;
  const int v = lo ? u : bs.N - u;
  const double2 a = elo[v & (kEmHi - 1)], b = ehi[v / kEmHi];
  const double2 r = make_double2(a.x + b.x + (a.x * b.x - a.y * b.y),
                                 a.y + b.y + (a.x * b.y + a.y * b.x));
  return make_double2(r.x, lo ? r.y : -r.y);
}

// E(x + y) = E(x) + E(y) + E(x) E(y) for E(t) = exp(-2 pi i t/N) - 1
__device__ __forceinline__ double2 em1_add(const double2 a, const double2 b) {
  return make_double2(a.x + b.x + (a.x * b.x - a.y * b.y), a.y + b.y + (a.x * b.y + a.y * b.x));
}

// sum_i g[i] (exp(-2 pi i k i/N) - 1) over the blocks [a_lo, a_hi) of 8
// consecutive terms: E(8ak + bk) = W^{8ak} E(bk) + E(8ak) exactly, so each
// block contributes (1 + E(8ak)) U_a + E(8ak) V_a, U_a = sum_b g E(bk),
// V_a = sum_b g.  Every term is formed without cancellation, so the error
// stays relative to the terms (small for small k), not to sum g.  E(8ak)
// starts from the table at a_lo and steps by E(8k) (em1_add): a run of at
// most a few steps adds a few roundings of |E(8ak)|.  8 k a < 2^31 for
// k <= N/2 and a < 2^11.
__device__ double2 head_sum(const BeamsetDev& bs, const double2* elo, const double2* ehi,
                            const double* g, const double* vsum, const int a_lo, const int a_hi,
                            const int k, const double2* eb, const double2 e8) {
  double2 acc = make_double2(0.0, 0.0);
  if (a_lo >= a_hi) return acc;
  double2 ea = em1_at(bs, elo, ehi, static_cast<uint32_t>(kHeadBlk) *
                                        static_cast<uint32_t>(k) * static_cast<uint32_t>(a_lo));
  for (int a = a_lo; a < a_hi; ++a) {
    const double* ga = g + kHeadBlk * a;
    double2 u = make_double2(0.0, 0.0);
#pragma unroll
    for (int b = 1; b < kHeadBlk; ++b) {
      u.x = fma(ga[b], eb[b].x, u.x);
      u.y = fma(ga[b], eb[b].y, u.y);
    }
    // (1 + ea) u + ea v
    const double2 t1 = make_double2(u.x + (ea.x * u.x - ea.y * u.y), u.y + (ea.x * u.y + ea.y * u.x));
    acc.x += t1.x + ea.x * vsum[a];
    acc.y += t1.y + ea.y * vsum[a];
    ea = em1_add(ea, e8);
  }
  return acc;
}

// The forward transform of one walker is split over `parts` workgroups
// (1, 2 or 4) by chunk groups: with one part the workgroup finishes the
// spectrum itself; with more, each writes its partial kept rows and partial
// moments and spectrum1d_kernel combines them by the group tree and
// finishes.  Either way the rows and moments are the same bits.
__global__ __launch_bounds__(kThreads) void forward1d_kernel(const BeamsetDev bs,
                                                              const WorkDev wk) {
  __shared__ XformSmem sm;
  const int tid = threadIdx.x;
  const int n = batch_n(wk);
  const int parts = batch_parts(n, wk.force_parts, wk.nslots, wk.parts_tab);
  const int walker = blockIdx.x / parts;
  const int part = blockIdx.x - walker * parts;
  if (walker >= n) return;
  double* st = walker_state(wk, walker);
  if (st[ST_STATUS] != POFD_OK) return;  // uniform across the workgroup
  unsigned long long tmark = wall_clock64();
  const int M2 = bs.M2;
  const double dflux = bs.dflux;
  const double sigmult = st[ST_SIGMULT];
  const double* coef = wk.coef + walker * wk.coef_stride;
  double2* zrows = wk.zrows + walker * wk.zrows_stride;
  double2* zdst = (parts == 1) ? zrows
                               : wk.zpart + static_cast<size_t>(blockIdx.x) * wk.zpart_stride;

  load_twiddles(bs, sm);
  if (tid < 32) sm.ex2[tid] = exp2(static_cast<double>(tid) * 0.03125) * dflux;
  {
    int kmaxb;
    const int nr = spectral_rows(bs, sigmult, sm.rowk2, sm.rowof, kmaxb);
    if (tid == 0) {
      sm.isc[0] = nr;
      sm.isc[1] = kmaxb;
      if (part == 0) {
        st[ST_KMAXB] = kmaxb;
        st[ST_NROWS] = nr;
      }
    }
  }
  __syncthreads();
  const int nrows = sm.isc[0];
  const int nchunks = M2 / kChunk;
  // this part's chunk groups (kGroups / parts of them) and their chunks
  const int g_lo = part * kGroups / parts, g_hi = (part + 1) * kGroups / parts;
  const int c_lo = group_chunk(g_lo, nchunks), c_hi = group_chunk(g_hi, nchunks);
  const int row4 = tid >> 2;
  // moments: this thread's share of the open group; the group tree in
  // sm.gtree
  double mn_part = 0.0, var_part = 0.0;
  tree_start(sm.gtree);
  // the rows' group tree lives in HBM: the open pair of G0, G1 in zdst and
  // of G2, G3 in zhi -- with one part the first noisy-row area, which
  // forward_tail fills only later
  double2* zhi = (parts == 1) ? zrows + bs.M : zdst;

  const ColOffsets co = col_offsets();
  for (int p0 = 0; p0 < nrows; p0 += kRowsPerPass) {
    const int np = min(kRowsPerPass, nrows - p0);
    // The pass's kept rows accumulate on the FP64 matrix cores:
    // [Re Z | Im Z] (256 x 16) += [Re Yt | Im Yt] (256 x 32) x
    // [[Re V, Im V], [-Im V, Re V]] (32 x 16) per chunk, Yt the twiddled
    // column transforms and V[n2][r] = W_M2^(n2 k2_r).  Wave wv takes rows
    // k1 in [16 wv, 16 wv + 16) in v_mfma_f64_16x16x4_f64 steps (lane l
    // holds A[l & 15][l >> 4], B[l >> 4][l & 15], D[i][l & 15] for
    // i = (l >> 4) + 4 rr in register rr): steps 2m and 2m + 1 take the
    // real and the imaginary parts of Yt at columns 4 m + (l >> 4); output
    // column j is Re (j < 8) or Im (j >= 8) of kept row p0 + (j & 7).
    f64x4 acc = {0.0, 0.0, 0.0, 0.0};
    const int lj = tid & 15, lk = (tid >> 4) & 3, k1w = 16 * (tid >> 6);
    const int lr = lj & 7;  // this lane's output row (of the pass)
    // closes, in order, every group whose chunks end at or before chunk
    // cend: its row sums and, in the first pass, its moments enter the
    // group tree; lanes j < 8 store (Re, Im) of row lr at rows k1 of
    // their registers, the imaginary parts from lane j + 8
    int gnext = g_lo;
    auto close_groups = [&](const int cend) {
      while (gnext < g_hi && group_chunk(gnext + 1, nchunks) <= cend) {
        const int g = gnext++;
        double2* zg = (g >= 2) ? zhi : zdst;
        const bool add = (g & 1) && g > g_lo;  // G_(g-1) is parked there
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const double im = __shfl_xor(acc[rr], 8);
          if (lj < 8 && lr < np) {
            double2* z = zg + (p0 + lr) * kM1 + k1w + lk + 4 * rr;
            const double2 v = make_double2(acc[rr], im);
            *z = add ? cadd(*z, v) : v;
          }
          acc[rr] = 0.0;
        }
        if (p0 == 0) {
          double a = mn_part, b = var_part;
          block_sum2(a, b, sm.red);
          tree_add(sm.gtree, a, b, g, parts);
          mn_part = 0.0;
          var_part = 0.0;
        }
      }
    };
    close_groups(c_lo);  // groups of no chunks ahead of this part's first
    // fill: thread (n1 = row4, q) makes R at 8 consecutive grid points of
    // chunk c into dst
    // the head of the grid (the first headn points, and the last headn with
    // a negative beam) enters the moments but not the transform:
    // forward_tail adds its part directly
    const bool head_row = (row4 == 0) || (bs.headneg && row4 == kM1 - 1);
    auto fill_chunk = [&](const int c, double2* dst) {
      // row and quad lane from the thread index through an empty asm, so
      // that the LDS offsets are recomputed per chunk rather than held
      int tf = threadIdx.x;
      asm volatile("" : "+v"(tf));
      const int row4 = tf >> 2, q = tf & 3;
      const int j0 = 2 * M2 * row4 + 2 * kChunk * c + 8 * q;
      // headn is a multiple of 8: a thread's 8 points are all in the head or
      // all outside it
      const bool hz = head_row && (j0 < bs.headn || j0 >= bs.N - bs.headn);
      // node intervals of the 8 fluxes (gsl_interp_accel_find, precomputed)
      const uint4 kq = *reinterpret_cast<const uint4*>(bs.kst + j0);
      const uint32_t kw[4] = {kq.x, kq.y, kq.z, kq.w};
      int k = kw[0] & 0xffff;
      double xk = bs.xnode[k];
      double4 cf = *reinterpret_cast<const double4*>(coef + 4 * k);
      double rprev = 0.0;
      // the common case: all 8 points in the positive fill range and below
      // wrapRidx - 1, where R needs no range test and the moments take the
      // plain terms (the same arithmetic as the general loop below)
      const int wr = bs.wrapRidx;
      const bool plain = j0 >= 1 && j0 >= bs.minit && j0 + 7 <= bs.maxit &&
                         (j0 + 7 < bs.minitn || j0 > bs.maxitn) && j0 + 7 < wr - 1;
      if (plain) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int j = j0 + e;
          const int ke = (kw[e >> 1] >> (16 * (e & 1))) & 0xffff;
          if (ke != k) {
            k = ke;
            xk = bs.xnode[k];
            cf = *reinterpret_cast<const double4*>(coef + 4 * k);
          }
          const double cflux = static_cast<double>(j) * dflux;
          const double delx = cflux - xk;
          const double sv = cf.x + delx * (cf.y + delx * (cf.z + delx * cf.w));
          const double rv = exp2_dflux(sv, sm.ex2);
          if (p0 == 0) {
            const double prod = cflux * rv;
            mn_part += prod;
            var_part += cflux * prod;
          }
          if (e & 1)
            dst[cpos(4 * q + (e >> 1), row4)] = hz ? make_double2(0.0, 0.0)
                                                   : make_double2(rprev, rv);
          rprev = rv;
        }
        return;
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int j = j0 + e;
        const int ke = (kw[e >> 1] >> (16 * (e & 1))) & 0xffff;
        if (ke != k) {
          k = ke;
          xk = bs.xnode[k];
          cf = *reinterpret_cast<const double4*>(coef + 4 * k);
        }
        double rv = 0.0;
        const bool inneg = (j >= bs.minitn && j <= bs.maxitn);
        if (inneg || (j >= bs.minit && j <= bs.maxit)) {
          // grid points above wrapRidx hold the wrapped negative fluxes
          const double cflux = inneg ? static_cast<double>(j) * dflux + bs.mval
                                     : static_cast<double>(j) * dflux;
          const double delx = cflux - xk;
          const double s = cf.x + delx * (cf.y + delx * (cf.z + delx * cf.w));
          rv = exp2_dflux(s, sm.ex2);
        }
        if (p0 == 0 && j >= 1) {
          // getMeanVarFromR (src/PDFactory.cc:625-679)
          double prod, cfl;
          if (j <= wr) {
            cfl = static_cast<double>(j) * dflux;
            prod = (j == wr) ? 0.5 * cfl * rv : cfl * rv;
          } else {
            cfl = static_cast<double>(j - wr - 1) * dflux + bs.minflux_R;
            prod = cfl * rv;
          }
          mn_part += prod;
          var_part += cfl * prod;
          if (j == wr - 1 && wr < bs.N - 1) {
            // the wrap correction reads R[wrapRidx - 1] (src/PDFactory.cc:
            // 661-668), kept as the reference has it
            const double cb = bs.minflux_R;
            const double pb = 0.5 * cb * rv;
            mn_part -= pb;
            var_part -= cb * pb;
          }
        }
        if (e & 1)
          dst[cpos(4 * q + (e >> 1), row4)] = hz ? make_double2(0.0, 0.0)
                                                 : make_double2(rprev, rv);
        rprev = rv;
      }
    };
    auto accum_chunk = [&](const int c, const double2* src) {
      // Yt at rows k1w + lj, columns 4 m + lk of the chunk, and V of those
      // columns at this lane's output row lr (zero past the pass's rows)
      // (one column block's real and imaginary steps after each other, so
      // that only its operands are live; the lane indices through an empty
      // asm, so that the LDS offsets are recomputed per chunk)
      int ta = threadIdx.x;
      asm volatile("" : "+v"(ta));
      const int lj = ta & 15, lk = (ta >> 4) & 3, k1w = 16 * (ta >> 6), lr = lj & 7;
      const int k2 = (lr < np) ? sm.rowk2[p0 + lr] : 0;
      const bool im_col = lj >= 8;
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int n2l = 4 * m + lk;
        const int n2 = kChunk * c + n2l;
        const int t = (k1w + lj) * n2;
        const double2 y =
            cmul(src[cpos(n2l, k1w + lj)], cmul(sm.wlo[t & (kM1 - 1)], sm.wm2[t >> 8]));
        const double2 v = (lr < np) ? sm.wm2[(n2 * k2) & (M2 - 1)] : make_double2(0.0, 0.0);
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(y.x, im_col ? v.y : v.x, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(y.y, im_col ? v.x : -v.y, acc, 0, 0, 0);
      }
    };
    // the two buffers alternate by chunk: a chunk's column transforms end
    // in the buffer it was filled into, so its row accumulation reads that
    // buffer while the next chunk fills the other, in one barrier interval
    // (the loop starts one chunk early, where it only fills the first chunk)
    for (int c = c_lo - 1; c < c_hi; ++c) {
      double2* cur = ((c - c_lo) & 1) ? sm.bufb : sm.bufa;
      double2* oth = ((c - c_lo) & 1) ? sm.bufa : sm.bufb;
      if (c >= c_lo) {
        fft256_cols<-1>(cur, oth, sm.wst, &co);
        accum_chunk(c, cur);
        close_groups(c + 1);
      }
      // each wave keeps the two phases in order (their registers do not
      // both fit); waves of the workgroup still overlap them
      __builtin_amdgcn_sched_barrier(0);
      if (c + 1 < c_hi) fill_chunk(c + 1, oth);
      __syncthreads();
    }
    // one part: (G0 + G1) + (G2 + G3); more parts leave their pair or
    // group in their slot
    if (parts == 1 && lj < 8 && lr < np) {
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int e = (p0 + lr) * kM1 + k1w + lk + 4 * rr;
        zdst[e] = cadd(zdst[e], zhi[e]);
      }
    }
  }
  if (parts > 1) {
    if (tid == 0) {
      wk.mpart[2 * blockIdx.x] = sm.gtree[0];
      wk.mpart[2 * blockIdx.x + 1] = sm.gtree[1];
    }
    return;
  }
  __syncthreads();  // thread 0's group tree
  const double mn = sm.gtree[2] + sm.gtree[0], var = sm.gtree[3] + sm.gtree[1];
  __threadfence_block();  // the row stores, visible to this workgroup
  __syncthreads();
  forward_tail(bs, wk, walker, sm, mn, var, tmark);
}

// the parts' kept rows and moments combined by the group tree, then the
// spectrum and the noisy rows (one workgroup per walker)
__global__ __launch_bounds__(kThreads) void spectrum1d_kernel(const BeamsetDev bs,
                                                               const WorkDev wk) {
  __shared__ XformSmem sm;
  const int tid = threadIdx.x;
  const int walker = blockIdx.x;
  const int n = batch_n(wk);
  const int parts = batch_parts(n, wk.force_parts, wk.nslots, wk.parts_tab);
  if (walker >= n || parts == 1) return;  // one part: the forward finished
  double* st = walker_state(wk, walker);
  if (st[ST_STATUS] != POFD_OK) return;  // uniform across the workgroup
  const unsigned long long tmark = wall_clock64();
  {
    int kmaxb;
    const int nr = spectral_rows(bs, st[ST_SIGMULT], sm.rowk2, sm.rowof, kmaxb);
    if (tid == 0) {
      sm.isc[0] = nr;
      sm.isc[1] = kmaxb;
    }
  }
  __syncthreads();
  const int nrows = sm.isc[0];
  // the group tree over the parts: 2 parts hold G0 + G1 and G2 + G3, 4
  // parts one group each
  const size_t p0 = static_cast<size_t>(walker) * parts;
  const double* mp = wk.mpart + 2 * p0;
  double mn, var;
  if (parts == 2) {
    mn = mp[0] + mp[2];
    var = mp[1] + mp[3];
  } else {
    mn = (mp[0] + mp[2]) + (mp[4] + mp[6]);
    var = (mp[1] + mp[3]) + (mp[5] + mp[7]);
  }
  double2* zrows = wk.zrows + walker * wk.zrows_stride;
  const double2* zp = wk.zpart + p0 * wk.zpart_stride;
  const size_t zs = wk.zpart_stride;
  for (int e = tid; e < nrows * kM1; e += kThreads) {
    zrows[e] = (parts == 2) ? cadd(zp[e], zp[zs + e])
                            : cadd(cadd(zp[e], zp[zs + e]), cadd(zp[2 * zs + e], zp[3 * zs + e]));
  }
  __threadfence_block();
  __syncthreads();
  forward_tail(bs, wk, walker, sm, mn, var, tmark);
}

// moments -> mean shift, R^(0), then the spectrum exp(R^ - R^0) with its
// phase and each dataset's noisy rows; zrows holds the kept rows and sm the
// row tables
__device__ void forward_tail(const BeamsetDev& bs, const WorkDev& wk, const int walker,
                             XformSmem& sm, const double mn, const double var,
                             unsigned long long tmark) {
  const int tid = threadIdx.x;
  const int M = bs.M;
  const double dflux = bs.dflux;
  double* st = walker_state(wk, walker);
  const double sigmult = st[ST_SIGMULT];
  double2* zrows = wk.zrows + walker * wk.zrows_stride;
  const int nrows = sm.isc[0], kmaxb = sm.isc[1];
  const double sgpos0 = sqrt(var);
  const bool doshift = (sgpos0 > dflux) && (fabs(mn) > dflux);
  const double shift = doshift ? -mn : 0.0;
  const double2 z0 = zrows[0];  // row k2 = 0 is always row position 0
  const double r0 = z0.x + z0.y;  // sum of R outside the head
  // the head of the grid (left out of the transform): R there, in the
  // free transform buffer -- hp[i] = R[i], i < H; hn[v] = R[N - v],
  // 1 <= v <= H (with a negative beam); V sums of kHeadBlk consecutive values
  const int H = bs.headn;
  const int nbp = H / kHeadBlk, nbn = bs.headneg ? H / kHeadBlk + 1 : 0;
  double* hp = reinterpret_cast<double*>(sm.bufa);
  double* hn = hp + H;                // H + kHeadBlk values, index v (hn[0] = 0)
  double* vp = hn + H + kHeadBlk;     // nbp
  double* vn = vp + nbp;              // nbn
  // the two twiddle tables of em1_at, in the other free transform buffer
  double2* elo = sm.bufb;             // kEmHi: E(l)
  double2* ehi = elo + kEmHi;         // M / kEmHi + 1: E(kEmHi h)
  {
    const double* coef = wk.coef + walker * wk.coef_stride;
    for (int i = tid; i < kEmHi; i += kThreads) elo[i] = bs.em1[i];
    for (int i = tid; i <= bs.M / kEmHi; i += kThreads) ehi[i] = bs.em1[i * kEmHi];
    if (tid < 32) sm.ex2[tid] = exp2(static_cast<double>(tid) * 0.03125) * dflux;
    __syncthreads();
    for (int i = tid; i < H; i += kThreads) hp[i] = r_grid(bs, coef, i, sm.ex2);
    for (int v = tid; v < H + kHeadBlk; v += kThreads)
      hn[v] = (bs.headneg && v >= 1 && v <= H) ? r_grid(bs, coef, bs.N - v, sm.ex2) : 0.0;
    __syncthreads();
    for (int a = tid; a < nbp + nbn; a += kThreads) {
      const double* g = (a < nbp) ? hp + kHeadBlk * a : hn + kHeadBlk * (a - nbp);
      double v = 0.0;
      for (int b = 0; b < kHeadBlk; ++b) v += g[b];
      ((a < nbp) ? vp : vn - nbp)[a] = v;
    }
    __syncthreads();
  }
  if (tid == 0) {
    double vh = 0.0;
    for (int a = 0; a < nbp; ++a) vh += vp[a];
    for (int a = 0; a < nbn; ++a) vh += vn[a];
    st[ST_MN] = mn;
    st[ST_VAR] = var;
    st[ST_DOSHIFT] = doshift ? 1.0 : 0.0;
    st[ST_SHIFT] = shift;
    st[ST_R0] = r0 + vh;
    st[ST_TFWD] = static_cast<double>(wall_clock64() - tmark);
    for (int f = 0; f < bs.nfields; ++f)
      sm.isc[2 + f] = spectral_kmax(sigmult * bs.fields[f].sigma_base, bs.iflux,
                                    bs.spectral_cut, M);
  }
  tmark = wall_clock64();

  // ---- r2c post-processing + prtrans = exp(R^ - R^0) e^{i(phase - shift w)}
  // (src/PDFactory.cc:1094-1132)
  double2* prtrans = wk.prtrans + walker * wk.spec_stride;
  // the head's part of R^(k) - R^(0), summed directly, parked in
  // prtrans[k]: kHeadSplit consecutive lanes share a k, each summing a
  // range of the blocks, and their partial sums are combined in lane order
  {
    constexpr int kHeadSplit = 4;
    const int nitems = (kmaxb + 1) * kHeadSplit;
    for (int it0 = 0; it0 < nitems; it0 += kThreads) {  // uniform trip count
      const int it = it0 + tid;
      const bool act = it < nitems;
      const int k = act ? it / kHeadSplit : 0;
      const int part = it % kHeadSplit;
      double2 dh = make_double2(0.0, 0.0);
      if (act) {
        // E(bk), b < 8, by steps of E(k) from the table; E(8k) for the blocks
        double2 eb[kHeadBlk];
        const double2 e1 = em1_at(bs, elo, ehi, static_cast<uint32_t>(k));
        eb[0] = make_double2(0.0, 0.0);
        eb[1] = e1;
#pragma unroll
        for (int b = 2; b < kHeadBlk; ++b) eb[b] = em1_add(eb[b - 1], e1);
        const double2 e8 = em1_add(eb[kHeadBlk - 1], e1);
        dh = head_sum(bs, elo, ehi, hp, vp, part * nbp / kHeadSplit,
                      (part + 1) * nbp / kHeadSplit, k, eb, e8);
        if (nbn) {
          const double2 dn = head_sum(bs, elo, ehi, hn, vn, part * nbn / kHeadSplit,
                                      (part + 1) * nbn / kHeadSplit, k, eb, e8);
          dh.x += dn.x;  // W^{(N - v) k} - 1 = conj(E(vk))
          dh.y -= dn.y;
        }
      }
#pragma unroll
      for (int off = 1; off < kHeadSplit; off <<= 1) {
        dh.x += __shfl_xor(dh.x, off);
        dh.y += __shfl_xor(dh.y, off);
      }
      if (act && part == 0) prtrans[k] = dh;
    }
  }
  __threadfence_block();  // other lanes read these head sums below
  __syncthreads();
  {
    const double wfac = shift * bs.iflux;
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
      const double2 dh = prtrans[k];  // the head sum, written above
      double expfac, ival;
      {
#pragma clang fp contract(off)
        expfac = exp((X.x - r0) + dh.x);
        ival = doshift ? (X.y + dh.y) - wfac * static_cast<double>(k) : X.y + dh.y;
      }
      double s2, c2;
      sincos(ival, &s2, &c2);
      prtrans[k] = make_double2(expfac * c2, expfac * s2);
    }
  }
  __threadfence_block();
  __syncthreads();

  // ---- per dataset: noise factor (src/PDFactory.cc:1172-1188) and the c2r
  // pre-processing into the kept rows, Zp = E + i O with
  // E = X[k] + conj(X[M-k]), O = (X[k] - conj(X[M-k])) e^{+i pi k/M};
  // the imaginary parts of X[0] and X[M] are ignored as FFTW's c2r does.
  for (int f = 0; f < bs.nfields; ++f) {
    const double sigma = sigmult * bs.fields[f].sigma_base;
    const double sigfac = -0.5 * sigma * sigma;
    const int kmaxf = sm.isc[2 + f];
    auto noisy = [&](int k) {
#pragma clang fp contract(off)
      if (k > kmaxf) return make_double2(0.0, 0.0);
      const double2 pt = prtrans[k];
      if (k == 0 || !(sigma > 0)) return pt;
      const double w = bs.iflux * static_cast<double>(k);
      const double ef = exp(sigfac * w * w);
      return make_double2(ef * pt.x, ef * pt.y);
    };
    double2* zp = zrows + static_cast<size_t>(1 + f) * M;
    for (int e = tid; e < nrows * kM1; e += kThreads) {
      const int r = e >> 8, k1 = e & (kM1 - 1);
      const int k = k1 + kM1 * sm.rowk2[r];
      double2 xa = noisy(k);
      double2 xb = noisy(M - k);
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
  }
  if (tid == 0) st[ST_TSPEC] = static_cast<double>(wall_clock64() - tmark);
}

// ===========================================================================
// inverse: one workgroup per (walker, dataset): the inverse transform from
// the noisy rows, clamp, P(D) and its 32-value block statistics to HBM,
// and the whole-array statistics of the split search
// ===========================================================================
__global__ __launch_bounds__(kThreads) void inverse1d_kernel(const BeamsetDev bs,
                                                              const WorkDev wk) {
  __shared__ XformSmem sm;
  const int tid = threadIdx.x;
  const int n = batch_n(wk);
  const int parts = batch_parts(n, wk.force_parts, wk.nslots, wk.parts_tab);
  const int walker = blockIdx.x / parts;
  const int part = blockIdx.x - walker * parts;
  const int f = blockIdx.y;
  if (walker >= n) return;
  const double* st = walker_state(wk, walker);
  if (st[ST_STATUS] != POFD_OK) return;  // uniform across the workgroup
  const unsigned long long t0 = wall_clock64();
  const int M = bs.M, M2 = bs.M2;
  const double dflux = bs.dflux;
  const FieldDev& fd = bs.fields[f];
  const size_t wf = static_cast<size_t>(walker) * bs.nfields + f;
  double* pofd = wk.pofd + wf * wk.pofd_stride;
  double* bmin = wk.bmin + wf * wk.blk_stride;
  double* bsum = wk.bsum + wf * wk.blk_stride;
  double* dbg = wk.dbg + wf * kDbg;
  const double2* zp =
      wk.zrows + walker * wk.zrows_stride + static_cast<size_t>(1 + f) * M;
  const double sigmult = st[ST_SIGMULT];
  const double var_noi = st[ST_VAR];
  const double sigma = sigmult * fd.sigma_base;
  // getPD: sgpos / sgneg (src/PDFactory.cc:1225-1232)
  const double sgpos = sqrt(var_noi + sigma * sigma);
  const double sgneg = (9 * sigma * sigma > var_noi) ? sigma : sgpos;

  load_twiddles(bs, sm);
  int kmaxb;
  const int nr_rows = spectral_rows(bs, sigmult, sm.rowk2, sm.rowof, kmaxb);
  if (tid == 0) {
    sm.isc[0] = nr_rows;
    // findSplitPoint search range (src/PDFactory.cc:790-801)
    const double n1sigpos = 2.5 * sgpos, n1signeg = 2.5 * sgneg;
    double bd = n1sigpos / dflux;
    bd = (bd > 2e9) ? 2e9 : bd;
    int topidx;
    if (n1sigpos > fabs(bs.minflux_R)) {
      topidx = bs.wrapRidx;
    } else {
      double td = -(n1signeg + bs.minflux_R) / dflux + bs.wrapRidx + 1;
      td = (td > 2e9) ? 2e9 : ((td < -2e9) ? -2e9 : td);
      topidx = static_cast<int>(td);
    }
    sm.isc[1] = static_cast<int>(bd);
    sm.isc[2] = topidx;
    if (part == 0) {
      dbg[7] = sgpos;
      dbg[8] = sgneg;
      dbg[12] = spectral_kmax(sigma, bs.iflux, bs.spectral_cut, M);
      dbg[13] = sm.isc[0];
      dbg[19] = kmaxb;
      dbg[23] = sigma;
    }
  }
  __syncthreads();
  const int nrows = sm.isc[0], botidx = sm.isc[1], topidx = sm.isc[2];
  const int nchunks = M2 / kChunk;
  const int g_lo = part * kGroups / parts, g_hi = (part + 1) * kGroups / parts;
  const int c_lo = group_chunk(g_lo, nchunks), c_hi = group_chunk(g_hi, nchunks);
  const int row4 = tid >> 2, q = tid & 3;

  // the sums of the P(D) and of j P(D): this thread's share of the open
  // chunk group; the group tree in sm.gtree
  double pk = -1.0, mnv = 0.0, s_part = 0.0, sj_part = 0.0;
  int mni = -1;
  tree_start(sm.gtree);
  int gnext = g_lo;
  auto close_groups = [&](const int cend) {
    while (gnext < g_hi && group_chunk(gnext + 1, nchunks) <= cend) {
      const int g = gnext++;
      double a = s_part, b = sj_part;
      block_sum2(a, b, sm.red);
      tree_add(sm.gtree, a, b, g, parts);
      s_part = 0.0;
      sj_part = 0.0;
    }
  };
  close_groups(c_lo);
  // The pruned transform across columns, B[k1][n2] = sum_r Z[k1][r]
  // conj(W_M2^(n2 k2_r)) for the chunk's 16 columns, as a real matrix
  // product on the FP64 matrix cores: [Re B | Im B] (256 x 32) =
  // [Re Z | Im Z] (256 x 2R) x [[Re V, -Im V], [Im V, Re V]] (2R x 32),
  // V[n2][r] = W_M2^(n2 k2_r), kept rows in blocks of R = 8.  Wave wv takes
  // rows k1 in [16 wv, 16 wv + 16) in v_mfma_f64_16x16x4_f64 steps (lane l
  // holds A[l & 15][l >> 4], B[l >> 4][l & 15], and D[i][l & 15] for
  // i = (l >> 4) + 4 rr in register rr): the four k slots of step s are the
  // real (s < 2) or imaginary parts of Z at rows r0 + (l >> 4) + 4 (s & 1).
  const int lj = tid & 15, lk = (tid >> 4) & 3, k1w = 16 * (tid >> 6);
  const bool one_block = nrows <= 8;
  double2 za0 = make_double2(0.0, 0.0), za1 = make_double2(0.0, 0.0);
  if (one_block) {  // the block's Z operands stay in registers
    if (lk < nrows) za0 = zp[lk * kM1 + k1w + lj];
    if (lk + 4 < nrows) za1 = zp[(lk + 4) * kM1 + k1w + lj];
  }
  for (int c = c_lo; c < c_hi; ++c) {
    {
      const int n2 = kChunk * c + lj;
      f64x4 dre = {0.0, 0.0, 0.0, 0.0}, dim = {0.0, 0.0, 0.0, 0.0};
      for (int r0 = 0; r0 < nrows; r0 += 8) {
        const int ra = r0 + lk, rb = r0 + lk + 4;
        double2 z0 = za0, z1 = za1;
        if (!one_block) {
          z0 = (ra < nrows) ? zp[ra * kM1 + k1w + lj] : make_double2(0.0, 0.0);
          z1 = (rb < nrows) ? zp[rb * kM1 + k1w + lj] : make_double2(0.0, 0.0);
        }
        const double2 v0 = (ra < nrows) ? sm.wm2[(n2 * sm.rowk2[ra]) & (M2 - 1)]
                                        : make_double2(0.0, 0.0);
        const double2 v1 = (rb < nrows) ? sm.wm2[(n2 * sm.rowk2[rb]) & (M2 - 1)]
                                        : make_double2(0.0, 0.0);
        dre = __builtin_amdgcn_mfma_f64_16x16x4f64(z0.x, v0.x, dre, 0, 0, 0);
        dim = __builtin_amdgcn_mfma_f64_16x16x4f64(z0.x, -v0.y, dim, 0, 0, 0);
        dre = __builtin_amdgcn_mfma_f64_16x16x4f64(z1.x, v1.x, dre, 0, 0, 0);
        dim = __builtin_amdgcn_mfma_f64_16x16x4f64(z1.x, -v1.y, dim, 0, 0, 0);
        dre = __builtin_amdgcn_mfma_f64_16x16x4f64(z0.y, v0.y, dre, 0, 0, 0);
        dim = __builtin_amdgcn_mfma_f64_16x16x4f64(z0.y, v0.x, dim, 0, 0, 0);
        dre = __builtin_amdgcn_mfma_f64_16x16x4f64(z1.y, v1.y, dre, 0, 0, 0);
        dim = __builtin_amdgcn_mfma_f64_16x16x4f64(z1.y, v1.x, dim, 0, 0, 0);
      }
      // the four-step twiddle, then column n2 of the chunk into LDS
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int k1 = k1w + lk + 4 * rr;
        const int t = k1 * n2;
        const double2 b = cmulc(make_double2(dre[rr], dim[rr]),
                                cmul(sm.wlo[t & (kM1 - 1)], sm.wm2[t >> 8]));
        sm.bufa[cpos(lj, k1)] = b;
      }
    }
    __syncthreads();
    fft256_cols<+1>(sm.bufa, sm.bufb, sm.wst);
    {
      const int n1 = row4;
      // lane q of the row's quad takes columns 4 cc + q, so that each of
      // its four 16-byte stores and the quad's other lanes' fill 64
      // contiguous bytes of the row
      const int jrow = 2 * M2 * n1 + 2 * kChunk * c;
      double v8[8];
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) {
        const double2 z = sm.bufa[cpos(4 * cc + q, n1)];
        v8[2 * cc] = z.x;
        v8[2 * cc + 1] = z.y;
      }
      double bm = 0.0, bsm = 0.0;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        double v = v8[e];
        if (v < 0) v = 0.0;  // unwrapAndNormalizePD positivity
        v8[e] = v;
        const int j = jrow + 8 * (e >> 1) + 2 * q + (e & 1);
        pk = (v > pk) ? v : pk;
        if (j >= botidx && j <= topidx && (mni < 0 || v < mnv)) {
          mnv = v;
          mni = j;
        }
        s_part += v;
        sj_part += static_cast<double>(j) * v;
        bm = (e == 0 || v < bm) ? v : bm;
        bsm += v;
      }
#pragma unroll
      for (int cc = 0; cc < 4; ++cc)
        *reinterpret_cast<double2*>(pofd + jrow + 8 * cc + 2 * q) =
            make_double2(v8[2 * cc], v8[2 * cc + 1]);
      // 32-value block statistics over the 4 lanes q = 0..3
      double o = __shfl_xor(bm, 1);
      bm = (o < bm) ? o : bm;
      o = __shfl_xor(bm, 2);
      bm = (o < bm) ? o : bm;
      bsm += __shfl_xor(bsm, 1);
      bsm += __shfl_xor(bsm, 2);
      if (q == 0) {
        // block jrow / kBlk = n1 * (M2 / kChunk) + c, stored at blk_slot
        bmin[c * kM1 + n1] = bm;
        bsum[c * kM1 + n1] = bsm;
      }
    }
    __syncthreads();
    close_groups(c + 1);
  }
  const double peak = block_max(pk, sm.red);
  block_argmin(mnv, mni, sm.red);
  if (tid == 0) {
    // one part: (G0 + G1) + (G2 + G3); more: this part's pair or group
    const double S = (parts == 1) ? sm.gtree[2] + sm.gtree[0] : sm.gtree[0];
    const double SJ = (parts == 1) ? sm.gtree[3] + sm.gtree[1] : sm.gtree[1];
    if (parts == 1) {
      dbg[1] = peak;
      dbg[2] = mni;  // -1: empty search range
      dbg[3] = mnv;
      dbg[20] = S;
      dbg[21] = SJ;
    } else {
      double* ip = wk.ipart + (static_cast<size_t>(blockIdx.x) * bs.nfields + f) * kIPart;
      ip[0] = peak;
      ip[1] = mni;
      ip[2] = mnv;
      ip[3] = S;
      ip[4] = SJ;
    }
    if (part == 0) dbg[28] = static_cast<double>(wall_clock64() - t0);
  }
}

// ===========================================================================
// finalize: one workgroup per (walker, dataset): findSplitPoint, rotation,
// normalisation, mean, binned log likelihood (an unbinned dataset gets
// log2 of its normalised P(D) in place for pixels1d_kernel)
// ===========================================================================
// combine for one walker: (curr - like_offset)/like_norm summed over the
// datasets of the beam set in order, added to the batch outputs (one
// thread; the walker's dataset results already written)
__device__ void combine_walker(const BeamsetDev& bs, const WorkDev& wk, const int walker,
                               const int accumulate, double* __restrict__ loglike,
                               int32_t* __restrict__ status) {
  const double* st = walker_state(wk, walker);
  const int s = static_cast<int>(st[ST_STATUS]);
  const double prev = accumulate ? loglike[walker] : 0.0;
  if (s == POFD_E_BADLIKE) {
    status[walker] = s;
    loglike[walker] = prev + 1e25;  // calcLikeSingle::bad_like per beam set
    return;
  }
  if (s != POFD_OK) {
    status[walker] = s;
    loglike[walker] = qnan();
    return;
  }
  double total = 0.0;
  for (int f = 0; f < bs.nfields; ++f) {
    const FieldDev& fd = bs.fields[f];
    const size_t wf = static_cast<size_t>(walker) * bs.nfields + f;
    double* dbg = wk.dbg + wf * kDbg;
    const int fs = static_cast<int>(dbg[15]);
    if (fs != POFD_OK) {  // the reference stops at the first failing dataset
      status[walker] = fs;
      loglike[walker] = qnan();
      return;
    }
    double curr = dbg[14];
    if (!fd.binned) {
      const double* up = wk.upart + wf * kPixBlocks;
      double s2 = 0.0;
      for (int b = 0; b < kPixBlocks; ++b) s2 += up[b];
      curr = 0.69314718055994530942 * s2;  // log2toe
      dbg[14] = curr;
    }
    total += (curr - fd.like_offset) / fd.like_norm;
  }
  status[walker] = POFD_OK;
  loglike[walker] = prev + total;
}

// fuse: the beam set has one dataset and it is binned, so the combine needs
// nothing from a later launch: thread 0 does it for the walker at the end
// (combine1d_kernel is then not launched)
__global__ __launch_bounds__(kThreads) void finalize1d_kernel(const BeamsetDev bs,
                                                               const WorkDev wk, const int fuse,
                                                               const int accumulate,
                                                               double* __restrict__ loglike,
                                                               int32_t* __restrict__ status_out) {
  __shared__ double red[kWaves * 2];
  __shared__ int isc[4];
  __shared__ double sc[2];
  const int tid = threadIdx.x;
  const int walker = blockIdx.x;
  const int f = blockIdx.y;
  const int n = batch_n(wk);
  const int parts = batch_parts(n, wk.force_parts, wk.nslots, wk.parts_tab);
  if (walker >= n) return;
  const double* st = walker_state(wk, walker);
  if (st[ST_STATUS] != POFD_OK) {  // uniform across the workgroup
    if (fuse && tid == 0) combine_walker(bs, wk, walker, accumulate, loglike, status_out);
    return;
  }
  unsigned long long tmark = wall_clock64();
  const int N = bs.N;
  const double dflux = bs.dflux;
  const FieldDev& fd = bs.fields[f];
  const size_t wf = static_cast<size_t>(walker) * bs.nfields + f;
  double* pofd = wk.pofd + wf * wk.pofd_stride;
  const double* bmin = wk.bmin + wf * wk.blk_stride;
  const double* bsum = wk.bsum + wf * wk.blk_stride;
  const int nbc = bs.M2 / kChunk;  // P(D) blocks per transform row
  double* dbg = wk.dbg + wf * kDbg;
  if (parts > 1) {
    // the inverse transform's parts: peak and argmin in part (= index)
    // order, the sums by the group tree (2 parts hold G0 + G1 and G2 + G3,
    // 4 parts one group each)
    if (tid == 0) {
      const double* ip0 =
          wk.ipart + (static_cast<size_t>(walker) * parts * bs.nfields + f) * kIPart;
      const size_t ps = static_cast<size_t>(bs.nfields) * kIPart;
      double pk = ip0[0], bv = ip0[2];
      int bi = static_cast<int>(ip0[1]);
      for (int p = 1; p < parts; ++p) {
        const double* ip = ip0 + p * ps;
        pk = (ip[0] > pk) ? ip[0] : pk;
        const int ui = static_cast<int>(ip[1]);
        const double u = ip[2];
        const bool take = (ui >= 0) && (bi < 0 || u < bv || (u == bv && ui < bi));
        bv = take ? u : bv;
        bi = take ? ui : bi;
      }
      double s, sj;
      if (parts == 2) {
        s = ip0[3] + ip0[ps + 3];
        sj = ip0[4] + ip0[ps + 4];
      } else {
        s = (ip0[3] + ip0[ps + 3]) + (ip0[2 * ps + 3] + ip0[3 * ps + 3]);
        sj = (ip0[4] + ip0[ps + 4]) + (ip0[2 * ps + 4] + ip0[3 * ps + 4]);
      }
      dbg[1] = pk;
      dbg[2] = bi;
      dbg[3] = bv;
      dbg[20] = s;
      dbg[21] = sj;
    }
    __threadfence_block();
    __syncthreads();
  }
  const double sgpos = dbg[7], sgneg = dbg[8];
  const double peak = dbg[1];
  const double S = dbg[20], SJ = dbg[21];
  int botidx, topidx;
  {
    const double n1sigpos = 2.5 * sgpos, n1signeg = 2.5 * sgneg;
    double bd = n1sigpos / dflux;
    bd = (bd > 2e9) ? 2e9 : bd;
    botidx = static_cast<int>(bd);
    if (n1sigpos > fabs(bs.minflux_R)) {
      topidx = bs.wrapRidx;
    } else {
      double td = -(n1signeg + bs.minflux_R) / dflux + bs.wrapRidx + 1;
      td = (td > 2e9) ? 2e9 : ((td < -2e9) ? -2e9 : td);
      topidx = static_cast<int>(td);
    }
  }

  // ---- findSplitPoint (src/PDFactory.cc:691-921) -----------------------
  if (tid == 0) {
    int status = POFD_OK;
    int split = -1;
    const double f1 = 5e-14, f2 = 1e-5, f3 = 5, n1 = 2.5;
    int minidx = static_cast<int>(dbg[2]);
    double minval = dbg[3];
    double jitter = 0.0;
    if (sgpos <= 0 || sgneg <= 0) status = POFD_E_SPLIT;
    const double fluxrange = N * dflux;
    if (status == POFD_OK && fluxrange < (n1 * sgpos + n1 * sgneg)) status = POFD_E_SPLIT;
    if (status == POFD_OK && peak <= 0.0) status = POFD_E_SPLIT;
    if (status == POFD_OK && (botidx >= N || topidx >= N || botidx < 0 || topidx < 0))
      status = POFD_E_SPLIT;
    if (status == POFD_OK && minidx < 0) {  // empty range: minidx = botidx
      minidx = botidx;
      minval = pofd[botidx];
    }
    if (status == POFD_OK && minval > f2 * peak) status = POFD_E_SPLIT;
    int decided = 0;
    if (status == POFD_OK) {
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
    if (status == POFD_OK && !decided) {
      if (minval <= 0.0) {
        const int nminup = 3;
        int bi = (minidx < botidx) ? botidx : minidx - nminup;
        const int ti = (minidx > topidx - nminup) ? topidx - nminup : minidx + nminup;
        if (bi < 0) bi = 0;
        for (int i = bi; i < ti && i < N; ++i) {
          const double cv = pofd[i];
          if (cv > minval) minval = cv;
        }
      }
      if (minval > 0) jitter = (f3 * minval) < jitter ? (f3 * minval) : jitter;
      if (botidx >= topidx)
        status = POFD_E_SPLIT;
      else if (pofd[topidx] < jitter)
        status = POFD_E_SPLIT;
    }
    isc[0] = status;
    isc[1] = split;
    isc[2] = decided;
    isc[3] = minidx;
    sc[0] = jitter;
    sc[1] = minval;
  }
  __syncthreads();
  if (isc[0] == POFD_OK && !isc[2]) {
    // largest i in [botidx, topidx-1] with pofd[i] < jitter; split = i+1
    const double jitter = sc[0];
    const int lo = botidx, hi = topidx - 1;
    const int bl = lo / kBlk, bh = hi / kBlk;
    int found;
    {  // top (possibly partial) block, value by value
      const int a = max(lo, bh * kBlk);
      int cand = -1;
      if (tid < kBlk) {
        const int i = a + tid;
        if (i <= hi && pofd[i] < jitter) cand = i;
      }
      found = block_max_int(cand, red);
    }
    if (found < 0 && bh - 1 >= bl + 1) {  // whole blocks via their minima
      for (int top = bh - 1; top >= bl + 1 && found < 0; top -= kThreads) {
        const int b = top - tid;
        int cb = -1;
        if (b >= bl + 1 && bmin[blk_slot(b, nbc)] < jitter) cb = b;
        const int bstar = block_max_int(cb, red);
        if (bstar >= 0) {
          int cand = -1;
          if (tid < kBlk) {
            const int i = bstar * kBlk + tid;
            if (pofd[i] < jitter) cand = i;
          }
          found = block_max_int(cand, red);
        }
      }
    }
    if (found < 0 && bl < bh) {  // bottom partial block
      int cand = -1;
      if (tid < kBlk) {
        const int i = lo + tid;
        if (i <= min(hi, bl * kBlk + kBlk - 1) && pofd[i] < jitter) cand = i;
      }
      found = block_max_int(cand, red);
    }
    if (tid == 0) {
      if (found < 0)
        isc[0] = POFD_E_SPLIT;
      else
        isc[1] = found + 1;
    }
  }
  __syncthreads();
  const unsigned long long t_split = wall_clock64() - tmark;
  tmark = wall_clock64();

  // ---- rotate, normalise, mean (unwrapAndNormalizePD) -------------------
  int status = isc[0];
  const int split = isc[1];
  double integral = 0.0, itot = 0.0, minflux = 0.0, Shigh = 0.0;
  if (status == POFD_OK) {
    // sum of the P(D) at indices >= split: whole blocks + the partial one
    const int bs0 = split / kBlk;
    double part = 0.0;
    for (int b = bs0 + 1 + tid; b < N / kBlk; b += kThreads) part += bsum[blk_slot(b, nbc)];
    if (tid < kBlk) {
      const int i = bs0 * kBlk + tid;
      if (i >= split && i < N) part += pofd[i];
    }
    Shigh = block_sum(part, red);
    const double plast = pofd[(split + N - 1) % N];
    const double pfirst = pofd[split];
    // trapezoid integral of the rotated array (PD::getIntegral)
    integral = (S - 0.5 * pfirst - 0.5 * plast) * dflux;
    itot = 1.0 / integral;
    // sum_i i pd[i] over the rotated array, last weight halved (PD::getMean)
    const double Mrot = SJ - static_cast<double>(split) * S +
                        static_cast<double>(N) * (S - Shigh);
    const double msum = (Mrot - 0.5 * static_cast<double>(N - 1) * plast) * itot;
    const double tmn = msum * (dflux * dflux);
    if (isinf(tmn) || isnan(tmn)) status = POFD_E_MEAN;
    minflux = -tmn;
  }

  // ---- log likelihood of a binned dataset (PD::getLogLikeBinned) --------
  double curr = 0.0;
  if (status == POFD_OK) {
#pragma clang fp contract(off)
    if (fd.binned) {
      const double maxflux = minflux + static_cast<double>(N - 1) * dflux;
      const double idflux = 1.0 / dflux;
      const double maxfluxval = maxflux - dflux;
      auto Lrot = [&](int i) {
        int jj = i + split;
        if (jj >= N) jj -= N;
        return log2_or_small(pofd[jj] * itot);
      };
      double part = 0.0;
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
      curr = 0.69314718055994530942 * block_sum(part, red);  // log2toe
    } else {
      __syncthreads();  // every load of the split search precedes the rewrite
      // eight values a step per thread (N is a multiple of 8 x kThreads):
      // their loads issued together
      for (int j = tid; j < N; j += 8 * kThreads) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = pofd[j + u * kThreads];
#pragma unroll
        for (int u = 0; u < 8; ++u) pofd[j + u * kThreads] = log2_or_small(v[u] * itot);
      }
    }
  }
  if (tid == 0) {
    dbg[0] = split;
    dbg[2] = isc[3];
    dbg[3] = sc[1];
    dbg[4] = sc[0];
    dbg[5] = integral;
    dbg[6] = -minflux;
    dbg[9] = st[ST_MN];
    dbg[10] = st[ST_VAR];
    dbg[11] = st[ST_R0];
    dbg[14] = curr;
    dbg[15] = status;
    dbg[16] = itot;
    dbg[17] = minflux;
    dbg[18] = (status == POFD_OK && !fd.binned) ? 1.0 : 0.0;
    dbg[22] = Shigh;
    dbg[24] = st[ST_TPREP];
    dbg[25] = 0.0;
    dbg[26] = st[ST_TFWD];
    dbg[27] = st[ST_TSPEC];
    dbg[29] = static_cast<double>(t_split);
    dbg[30] = static_cast<double>(wall_clock64() - tmark);
    dbg[31] = dbg[24] + dbg[26] + dbg[27] + dbg[28] + dbg[29] + dbg[30];
    if (fuse) combine_walker(bs, wk, walker, accumulate, loglike, status_out);
  }
}

// ===========================================================================
// pixels: unbinned log likelihood, kPixBlocks workgroups per (walker, field)
// ===========================================================================
constexpr int kPixThreads = 256;
constexpr int kPixU = 8;  // pixels per thread per step

__global__ __launch_bounds__(kPixThreads, 8) void pixels1d_kernel(const BeamsetDev bs,
                                                                const WorkDev wk) {
  __shared__ double red[kPixThreads / 64];
  const int blk = blockIdx.x;
  const int walker = blockIdx.y;
  const int f = blockIdx.z;
  const FieldDev& fd = bs.fields[f];
  if (fd.binned || walker >= batch_n(wk)) return;
  const size_t wf = static_cast<size_t>(walker) * bs.nfields + f;
  double* out = wk.upart + wf * kPixBlocks + blk;
  const double* dbg = wk.dbg + wf * kDbg;
  if (walker_state(wk, walker)[ST_STATUS] != POFD_OK || dbg[15] != POFD_OK ||
      dbg[18] == 0.0) {
    if (threadIdx.x == 0) *out = 0.0;
    return;
  }
  const int N = bs.N;
  const double dflux = bs.dflux;
  const int split = static_cast<int>(dbg[0]);
  const double minflux = dbg[17];
  const double* L = wk.pofd + wf * wk.pofd_stride;  // log2 P(D), unrotated
  double part = 0.0;
  {
#pragma clang fp contract(off)
    const double maxflux = minflux + static_cast<double>(N - 1) * dflux;
    const double idflux = 1.0 / dflux;
    const double maxfluxval = maxflux - dflux;
    // rotated index i -> i + split mod N (N is a power of two), as a 32-bit
    // byte offset from the walker's table
    const char* Lb = reinterpret_cast<const char*>(L);
    const uint32_t nm1 = static_cast<uint32_t>(N - 1), usplit = static_cast<uint32_t>(split);
    auto Lr = [&](const uint32_t i) {
      return *reinterpret_cast<const double*>(Lb + (((i + usplit) & nm1) << 3));
    };
    // contiguous pixel ranges per workgroup: pixels are sorted, so
    // neighbouring lanes read neighbouring P(D) values
    const uint64_t per = (fd.npix + kPixBlocks - 1) / kPixBlocks;
    const uint64_t p0 = per * blk;
    const uint64_t p1 = (p0 + per < fd.npix) ? p0 + per : fd.npix;
    const uint32_t cnt = (p1 > p0) ? static_cast<uint32_t>(p1 - p0) : 0u;
    const char* pb = reinterpret_cast<const char*>(fd.pix + p0);
    auto pix = [&](const uint32_t k) { return *reinterpret_cast<const double*>(pb + (k << 3)); };
    // the term of pixel flux cflux, as PD::getLogLike takes it: Lr(0) at or
    // below minflux, Lr(N - 1) at or above maxflux - dflux, else the linear
    // interpolation; the end cases as weights (1, 0) of the same formula,
    // which give the end value exactly (the table holds finite values)
    auto term = [&](const double cflux) {
      const bool lo = cflux <= minflux, hi = cflux >= maxfluxval;
      const double delt = (lo || hi) ? 0.0 : (cflux - minflux) * idflux;
      const uint32_t id0 = static_cast<uint32_t>(static_cast<int>(delt));
      const double t0 = delt - static_cast<double>(static_cast<int>(id0));
      const uint32_t idx = hi ? nm1 : id0;
      const double t = hi ? 0.0 : t0;
      return (1.0 - t) * Lr(idx) + t * Lr(idx + 1);
    };
    // kPixU pixels a step: their gathers are issued together, and the next
    // step's pixel loads before them, so a step waits for one memory
    // latency, not two; the terms are added in pixel order as a one-pixel
    // loop would
    constexpr uint32_t kStep = static_cast<uint32_t>(kPixU) * kPixThreads;
    constexpr uint32_t kLast = static_cast<uint32_t>(kPixU - 1) * kPixThreads;
    uint32_t i = threadIdx.x;
    double cf[kPixU];
    if (i + kLast < cnt) {
#pragma unroll
      for (int u = 0; u < kPixU; ++u) cf[u] = pix(i + u * kPixThreads);
    }
    for (; i + kLast < cnt; i += kStep) {
      double nx[kPixU];
      const bool more = i + kStep + kLast < cnt;  // the loop's next condition
#pragma unroll
      for (int u = 0; u < kPixU; ++u) nx[u] = more ? pix(i + kStep + u * kPixThreads) : 0.0;
      double tv[kPixU];
#pragma unroll
      for (int u = 0; u < kPixU; ++u) tv[u] = term(cf[u]);
#pragma unroll
      for (int u = 0; u < kPixU; ++u) part += tv[u];
#pragma unroll
      for (int u = 0; u < kPixU; ++u) cf[u] = nx[u];
    }
    for (; i < cnt; i += kPixThreads) part += term(pix(i));
  }
  const double s = block_sum(part, red);
  if (threadIdx.x == 0) *out = s;
}

// ===========================================================================
// combine: per walker, (curr - like_offset)/like_norm summed over the
// datasets of the beam set in order, added to the batch outputs
// ===========================================================================
__global__ void combine1d_kernel(const BeamsetDev bs, const WorkDev wk,
                                 const int accumulate, double* __restrict__ loglike,
                                 int32_t* __restrict__ status) {
  const int walker = blockIdx.x * blockDim.x + threadIdx.x;
  if (walker >= batch_n(wk)) return;
  combine_walker(bs, wk, walker, accumulate, loglike, status);
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

int launch_loglike1d(const BeamsetDev& bs, const WorkDev& wk_in, int beamset_idx,
                     const float* d_theta, uint32_t nwalk, const int32_t* d_count,
                     uint32_t stride, double* d_loglike, int32_t* d_status,
                     int accumulate, int force_parts, const uint8_t* parts_tab_host,
                     const uint8_t* parts_tab_dev, void* stream, hipEvent_t* ev) {
  if (force_parts < 0 || force_parts > 4) return POFD_E_ARG;
  WorkDev wk = wk_in;
  wk.dcount = d_count;
  wk.nhost = static_cast<int>(nwalk);
  wk.force_parts = force_parts;
  wk.parts_tab = parts_tab_dev;
  // workgroups of the split transforms: exactly n x parts for a host count;
  // for a device count every batch up to nwalk fits in max(nwalk, nslots)
  uint32_t nxf;
  bool split;
  if (d_count) {
    nxf = (nwalk > wk.nslots) ? nwalk : wk.nslots;
    split = true;
  } else {
    const int parts = batch_parts(static_cast<int>(nwalk), force_parts, wk.nslots,
                                  parts_tab_host);
    nxf = nwalk * static_cast<uint32_t>(parts);
    split = parts > 1;
  }
  (void)beamset_idx;
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto mark = [&](int i) {
    if (ev) (void)hipEventRecord(ev[i], s);
  };
  mark(POFD_STAGE_PREP);
  hipLaunchKernelGGL(prep1d_kernel, dim3(split ? nxf : nwalk), dim3(kThreads), 0, s, bs, wk,
                     d_theta, stride, d_status, accumulate);
  if (split)
    hipLaunchKernelGGL(prepspline1d_kernel, dim3(nwalk), dim3(kThreads), 0, s, bs, wk);
  mark(POFD_STAGE_FORWARD);
  hipLaunchKernelGGL(forward1d_kernel, dim3(nxf), dim3(kThreads), 0, s, bs, wk);
  if (split)
    hipLaunchKernelGGL(spectrum1d_kernel, dim3(nwalk), dim3(kThreads), 0, s, bs, wk);
  mark(POFD_STAGE_INVERSE);
  hipLaunchKernelGGL(inverse1d_kernel, dim3(nxf, bs.nfields), dim3(kThreads), 0, s, bs, wk);
  bool any_unbinned = false;
  for (int f = 0; f < bs.nfields; ++f) any_unbinned |= !bs.fields[f].binned;
  // one binned dataset: finalize does the combine too
  const int fuse = (bs.nfields == 1 && !any_unbinned) ? 1 : 0;
  mark(POFD_STAGE_FINALIZE);
  hipLaunchKernelGGL(finalize1d_kernel, dim3(nwalk, bs.nfields), dim3(kThreads), 0,
                     s, bs, wk, fuse, accumulate, d_loglike, d_status);
  mark(POFD_STAGE_PIXELS);
  if (any_unbinned)
    hipLaunchKernelGGL(pixels1d_kernel, dim3(kPixBlocks, nwalk, bs.nfields),
                       dim3(kPixThreads), 0, s, bs, wk);
  mark(POFD_STAGE_COMBINE);
  if (!fuse)
    hipLaunchKernelGGL(combine1d_kernel, dim3((nwalk + 63) / 64), dim3(64), 0, s,
                       bs, wk, accumulate, d_loglike, d_status);
  mark(POFD_NSTAGES);
  return hipGetLastError() == hipSuccess ? POFD_OK : POFD_E_INTERNAL;
}

// calcLike::getLogLike's prior terms (src/calcLike.cc:923-952): Gaussian
// priors on the sigma multiplier (mean 1), on the mean flux per area (the
// CFIRB) and on the mean flux^2 per area (the Poisson noise), and the
// difference regularisation of the knot slopes in log2 space
// (numberCountsKnotsSpline::differenceRegularize, src/numberCountsKnotsSpline.cc:540-559);
// walkers whose datasets failed keep their status and NaN
__global__ void prior1d_kernel(const BeamsetDev bs, const WorkDev wk, const PriorDev pr,
                               const float* __restrict__ theta, const uint32_t stride,
                               double* __restrict__ loglike, const int32_t* __restrict__ status) {
  const int walker = blockIdx.x * blockDim.x + threadIdx.x;
  if (walker >= batch_n(wk)) return;
  const int s = status[walker];
  if (s != POFD_OK && s != POFD_E_BADLIKE) return;
  const double h = 0.918938517570495605469;  // log(2 pi) / 2
  const float* t = theta + static_cast<size_t>(walker) * stride;
  const double* bo = wk.bonus + static_cast<size_t>(walker) * 4;
  double add = 0.0;
  if (pr.flags & kPriorSigma) {
    const double v = (static_cast<double>(t[bs.nknots]) - 1.0) / pr.sigma_width;
    add -= h + log(pr.sigma_width) + 0.5 * v * v;
  }
  if (pr.flags & kPriorCfirb) {
    const double v = (pr.cfirb_mean - bo[1]) / pr.cfirb_sigma;
    add -= h + log(pr.cfirb_sigma) + 0.5 * v * v;
  }
  if (pr.flags & kPriorPoisson) {
    const double v = (pr.poisson_mean - bo[2]) / pr.poisson_sigma;
    add -= h + log(pr.poisson_sigma) + 0.5 * v * v;
  }
  if (pr.reg_alpha > 0.0 && bs.nknots > 1) {
    const int n = bs.nknots;
    double mean = 0.0;
    for (int i = 1; i < n; ++i)
      mean += (bs.logfac * static_cast<double>(t[i]) - bs.logfac * static_cast<double>(t[i - 1])) /
              (bs.logknots[i] - bs.logknots[i - 1]);
    mean /= static_cast<double>(n - 1);
    double pen = 0.0;
    for (int i = 1; i < n; ++i) {
      const double d = (bs.logfac * static_cast<double>(t[i]) -
                        bs.logfac * static_cast<double>(t[i - 1])) /
                           (bs.logknots[i] - bs.logknots[i - 1]) -
                       mean;
      pen -= d * d;
    }
    add += pr.reg_alpha * pen;
  }
  loglike[walker] += add;
}

int launch_prior1d(const BeamsetDev& bs0, const WorkDev& wk0, const PriorDev& pr,
                   const float* d_theta, uint32_t stride, uint32_t nwalk, const int32_t* d_count,
                   double* d_loglike, const int32_t* d_status, void* stream) {
  WorkDev wk = wk0;
  wk.dcount = d_count;
  wk.nhost = static_cast<int>(nwalk);
  hipLaunchKernelGGL(prior1d_kernel, dim3((nwalk + 63) / 64), dim3(64), 0,
                     static_cast<hipStream_t>(stream), bs0, wk, pr, d_theta, stride, d_loglike,
                     d_status);
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
