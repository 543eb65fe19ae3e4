// PD::getLogLike and PDDouble::getLogLike on the device: the interpolated
// log likelihood of a dataset given a P(D) the caller holds
// (src/PD.cc:538-677, src/PDDouble.cc:976-1222).
//
// The likelihood batch computes these sums inside its own kernels; these
// are the stand-alone entry points behind the reference's PD / PDDouble
// API.  One thread per pixel (or histogram bin), grid-stride; each
// workgroup reduces in LDS to one partial sum and a single workgroup adds
// the partials in index order, so the result does not depend on timing.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace pofd {
namespace {

constexpr int kPdlThreads = 256;
constexpr int kPdlBlocks = 1024;  // partial sums (grid-stride beyond that)

struct Pd1Args {
  const double* pd;
  int n;
  double minflux, dflux, idflux, maxfluxval;
  int islog;
  const double* pix;      // unbinned
  uint64_t npix;
  const uint32_t* bins;   // binned
  uint32_t nbins;
  double bincent0, bindelta;
};

struct Pd2Args {
  const double* pd;       // n1 x n2, band-1 index outer
  int n1, n2;
  double minflux1, minflux2, idflux1, idflux2, maxflux1, maxflux2;
  int islog;
  const double* pix1;
  const double* pix2;
  uint64_t npix;
  const uint32_t* bins;   // nb1 x nb2
  uint32_t nb1, nb2;
  double bincent01, bindelta1, bincent02, bindelta2;
};

__device__ inline double lg(double v, int islog) { return islog ? v : log2(v); }

// PD::getLogLikeUnbinned (src/PD.cc:628-677): edges x <= minflux and
// x >= maxflux - dflux
__device__ double term1_unbinned(const Pd1Args& a, double x) {
  if (x <= a.minflux) return lg(a.pd[0], a.islog);
  if (x >= a.maxfluxval) return lg(a.pd[a.n - 1], a.islog);
  const double delt = (x - a.minflux) * a.idflux;
  const int idx = static_cast<int>(delt);
  const double t = delt - static_cast<double>(idx);
  return (1.0 - t) * lg(a.pd[idx], a.islog) + t * lg(a.pd[idx + 1], a.islog);
}

// PD::getLogLikeBinned (src/PD.cc:557-619): edges c < minflux and
// c > maxflux - dflux, weighted by the bin count
__device__ double term1_binned(const Pd1Args& a, uint32_t i) {
  const uint32_t nin = a.bins[i];
  if (nin == 0) return 0.0;
  const double c = a.bincent0 + static_cast<double>(i) * a.bindelta;
  double v;
  if (c < a.minflux) {
    v = lg(a.pd[0], a.islog);
  } else if (c > a.maxfluxval) {
    v = lg(a.pd[a.n - 1], a.islog);
  } else {
    const double delt = (c - a.minflux) * a.idflux;
    const int idx = static_cast<int>(delt);
    const double t = delt - static_cast<double>(idx);
    v = (1.0 - t) * lg(a.pd[idx], a.islog) + t * lg(a.pd[idx + 1], a.islog);
  }
  return v * static_cast<double>(nin);
}

// PDDouble interpolation for one (f1, f2).  `strict` selects the edge tests:
// the log-stored unbinned sum tests <= / >= (src/PDDouble.cc:1011-1045),
// the linear-stored unbinned sum and both binned sums test < / >
// (:1047-1079, :1136-1216).  Edges are against maxflux, not maxflux - dflux.
__device__ double interp2(const Pd2Args& a, double c1, double c2, bool strict) {
  const double d1 = (c1 - a.minflux1) * a.idflux1;
  const double d2 = (c2 - a.minflux2) * a.idflux2;
  const int i1 = static_cast<int>(d1);
  const int i2 = static_cast<int>(d2);
  const int n1 = a.n1, n2 = a.n2;
  const size_t n2n1 = static_cast<size_t>(n2) * n1;
  const size_t n2i1 = static_cast<size_t>(n2) * i1;
  const bool lo1 = strict ? (c1 < a.minflux1) : (c1 <= a.minflux1);
  const bool hi1 = strict ? (c1 > a.maxflux1) : (c1 >= a.maxflux1);
  const bool lo2 = strict ? (c2 < a.minflux2) : (c2 <= a.minflux2);
  const bool hi2 = strict ? (c2 > a.maxflux2) : (c2 >= a.maxflux2);
  const double* pd = a.pd;
  const int il = a.islog;
  if (lo1) {
    if (lo2) return lg(pd[0], il);
    if (hi2) return lg(pd[n2 - 1], il);
    return lg(pd[i2], il);
  }
  if (hi1) {
    if (lo2) return lg(pd[n2n1 - n1], il);
    if (hi2) return lg(pd[n2n1 - 1], il);
    return lg(pd[n2n1 - n1 + i2], il);
  }
  if (lo2) return lg(pd[n2i1], il);
  if (hi2) return lg(pd[n2i1 + n2 - 1], il);
  const double t = d1 - static_cast<double>(i1);
  const double u = d2 - static_cast<double>(i2);
  const double omu = 1.0 - u, omt = 1.0 - t;
  const size_t b = n2i1 + i2;
  return omt * (omu * lg(pd[b], il) + u * lg(pd[b + 1], il)) +
         t * (omu * lg(pd[b + n2], il) + u * lg(pd[b + n2 + 1], il));
}

__device__ double block_reduce(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kPdlThreads / 2; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  return red[0];
}

__global__ __launch_bounds__(kPdlThreads) void pd1_like_kernel(const Pd1Args a,
                                                               double* __restrict__ part) {
  __shared__ double red[kPdlThreads];
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kPdlThreads;
  double s = 0.0;
  if (a.bins) {
    for (uint64_t i = blockIdx.x * kPdlThreads + threadIdx.x; i < a.nbins; i += stride)
      s += term1_binned(a, static_cast<uint32_t>(i));
  } else {
    for (uint64_t i = blockIdx.x * kPdlThreads + threadIdx.x; i < a.npix; i += stride)
      s += term1_unbinned(a, a.pix[i]);
  }
  const double t = block_reduce(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(kPdlThreads) void pd2_like_kernel(const Pd2Args a,
                                                               double* __restrict__ part) {
  __shared__ double red[kPdlThreads];
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kPdlThreads;
  double s = 0.0;
  if (a.bins) {
    const uint64_t nb = static_cast<uint64_t>(a.nb1) * a.nb2;
    for (uint64_t k = blockIdx.x * kPdlThreads + threadIdx.x; k < nb; k += stride) {
      const uint32_t nin = a.bins[k];
      if (nin == 0) continue;
      const uint32_t i = static_cast<uint32_t>(k / a.nb2), j = static_cast<uint32_t>(k % a.nb2);
      const double c1 = a.bincent01 + static_cast<double>(i) * a.bindelta1;
      const double c2 = a.bincent02 + static_cast<double>(j) * a.bindelta2;
      s += static_cast<double>(nin) * interp2(a, c1, c2, true);
    }
  } else {
    const bool strict = a.islog == 0;
    for (uint64_t k = blockIdx.x * kPdlThreads + threadIdx.x; k < a.npix; k += stride)
      s += interp2(a, a.pix1[k], a.pix2[k], strict);
  }
  const double t = block_reduce(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// the partials in index order, times ln 2 (pofd_mcmc::log2toe)
__global__ __launch_bounds__(kPdlThreads) void pd_like_sum_kernel(const double* __restrict__ part,
                                                                  int nblk,
                                                                  double* __restrict__ out) {
  __shared__ double red[kPdlThreads];
  double s = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kPdlThreads) s += part[i];
  const double t = block_reduce(s, red);
  if (threadIdx.x == 0) out[0] = 0.69314718055994530942 * t;
}

int blocks_for(uint64_t n) {
  const uint64_t b = (n + kPdlThreads - 1) / kPdlThreads;
  return static_cast<int>(b < 1 ? 1 : (b > static_cast<uint64_t>(kPdlBlocks) ? kPdlBlocks : b));
}

}  // namespace

size_t pd_like_scratch_doubles() { return kPdlBlocks + 1; }

int launch_pd_loglike1d(const double* d_pd, int n, double minflux, double dflux, int islog,
                        const double* d_pix, uint64_t npix, const uint32_t* d_bins,
                        uint32_t nbins, double bincent0, double bindelta, double* d_scratch,
                        void* stream) {
  Pd1Args a{};
  a.pd = d_pd;
  a.n = n;
  a.minflux = minflux;
  a.dflux = dflux;
  a.idflux = 1.0 / dflux;
  a.maxfluxval = (minflux + static_cast<double>(n - 1) * dflux) - dflux;
  a.islog = islog;
  a.pix = d_pix;
  a.npix = npix;
  a.bins = d_bins;
  a.nbins = nbins;
  a.bincent0 = bincent0;
  a.bindelta = bindelta;
  const int nb = blocks_for(d_bins ? nbins : npix);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(pd1_like_kernel, dim3(nb), dim3(kPdlThreads), 0, s, a, d_scratch);
  hipLaunchKernelGGL(pd_like_sum_kernel, dim3(1), dim3(kPdlThreads), 0, s, d_scratch, nb,
                     d_scratch + kPdlBlocks);
  return hipGetLastError() == hipSuccess ? POFD_OK : POFD_E_INTERNAL;
}

int launch_pd_loglike2d(const double* d_pd, int n1, int n2, double minflux1, double dflux1,
                        double minflux2, double dflux2, int islog, const double* d_pix1,
                        const double* d_pix2, uint64_t npix, const uint32_t* d_bins,
                        uint32_t nb1, uint32_t nb2, double bincent01, double bindelta1,
                        double bincent02, double bindelta2, double* d_scratch, void* stream) {
  Pd2Args a{};
  a.pd = d_pd;
  a.n1 = n1;
  a.n2 = n2;
  a.minflux1 = minflux1;
  a.minflux2 = minflux2;
  a.idflux1 = 1.0 / dflux1;
  a.idflux2 = 1.0 / dflux2;
  a.maxflux1 = minflux1 + static_cast<double>(n1 - 1) * dflux1;
  a.maxflux2 = minflux2 + static_cast<double>(n2 - 1) * dflux2;
  a.islog = islog;
  a.pix1 = d_pix1;
  a.pix2 = d_pix2;
  a.npix = npix;
  a.bins = d_bins;
  a.nb1 = nb1;
  a.nb2 = nb2;
  a.bincent01 = bincent01;
  a.bindelta1 = bindelta1;
  a.bincent02 = bincent02;
  a.bindelta2 = bindelta2;
  const int nb = blocks_for(d_bins ? static_cast<uint64_t>(nb1) * nb2 : npix);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(pd2_like_kernel, dim3(nb), dim3(kPdlThreads), 0, s, a, d_scratch);
  hipLaunchKernelGGL(pd_like_sum_kernel, dim3(1), dim3(kPdlThreads), 0, s, d_scratch, nb,
                     d_scratch + kPdlBlocks);
  return hipGetLastError() == hipSuccess ? POFD_OK : POFD_E_INTERNAL;
}

}  // namespace pofd
