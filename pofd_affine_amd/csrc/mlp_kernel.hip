// Kernels of the log-likelihood surrogate (a small MLP): a tiled GEMM on the
// matrix cores with bfloat16 or float operands and float arithmetic, GELU
// forward and backward, the squared-error gradient, an Adam step, and the
// fused inference of the whole network.
#include <hip/hip_runtime.h>

#include "mlp_infer.h"
#include "mlp_internal.h"

namespace pofd_mlp_dev {

namespace {

template <int T>
__device__ __forceinline__ float ld(const void* p, size_t i) {
  if (T == 0) return bf2f(static_cast<const bf16*>(p)[i]);
  return static_cast<const float*>(p)[i];
}

constexpr int kTile = 64, kKs = 32;

// One K-slice of A and B into registers (kKs x 64 of each per workgroup,
// 8 values per thread); zeros outside the matrices.
template <int TA, int TB>
__device__ __forceinline__ void load_slice(const void* A, int lda, int ta, const void* B, int ldb,
                                           int tb, int M, int N, int K, int m0, int n0, int k0,
                                           float* va, float* vb) {
  const int tid = threadIdx.x;
  const bool inner = (m0 + kTile <= M) && (n0 + kTile <= N) && (k0 + kKs <= K);
#pragma unroll
  for (int q = 0; q < kKs * kTile / 256; ++q) {
    const int e = tid + 256 * q;
    const int kk = e / kTile, mm = e % kTile;
    const int m = m0 + mm, n = n0 + mm, k = k0 + kk;
    if (inner) {  // uniform over the workgroup: unguarded loads
      va[q] = ld<TA>(A, ta ? static_cast<size_t>(k) * lda + m : static_cast<size_t>(m) * lda + k);
      vb[q] = ld<TB>(B, tb ? static_cast<size_t>(n) * ldb + k : static_cast<size_t>(k) * ldb + n);
    } else {
      va[q] = (m < M && k < K)
                  ? ld<TA>(A, ta ? static_cast<size_t>(k) * lda + m : static_cast<size_t>(m) * lda + k)
                  : 0.0f;
      vb[q] = (n < N && k < K)
                  ? ld<TB>(B, tb ? static_cast<size_t>(n) * ldb + k : static_cast<size_t>(k) * ldb + n)
                  : 0.0f;
    }
  }
}

// C = A B (+ bias) on the matrix cores in float: a 64 x 64 tile per
// workgroup of four waves, each wave 2 x 2 tiles of v_mfma_f32_16x16x4_f32
// (lane l holds A[l%16][l/16] and B[l/16][l%16] of a 16 x 4 / 4 x 16 step
// and receives D[4(l/16)..+3][l%16]).  The next K-slice is loaded into
// registers while the current one is multiplied out of LDS.
typedef float float4m __attribute__((ext_vector_type(4)));

template <int TA, int TB>
__global__ __launch_bounds__(256) void gemm_kernel(const void* A, int lda, int ta, const void* B,
                                                   int ldb, int tb, float* C, int ldc,
                                                   const float* bias, int M, int N, int K) {
  __shared__ float As[kKs][kTile + 4];
  __shared__ float Bs[kKs][kTile + 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  float4m acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = float4m{0.0f, 0.0f, 0.0f, 0.0f};
  constexpr int kPer = kKs * kTile / 256;
  float va[kPer], vb[kPer];
  load_slice<TA, TB>(A, lda, ta, B, ldb, tb, M, N, K, m0, n0, 0, va, vb);
  for (int k0 = 0; k0 < K; k0 += kKs) {
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int e = tid + 256 * q;
      As[e / kTile][e % kTile] = va[q];
      Bs[e / kTile][e % kTile] = vb[q];
    }
    __syncthreads();
    if (k0 + kKs < K)
      load_slice<TA, TB>(A, lda, ta, B, ldb, tb, M, N, K, m0, n0, k0 + kKs, va, vb);
#pragma unroll
    for (int kk = 0; kk < kKs; kk += 4) {
      float a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[kk + lk][wm + 16 * i + li];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Bs[kk + lk][wn + 16 * j + li];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + wn + 16 * j + li;
      if (n >= N) continue;
      const float bn = bias ? bias[n] : 0.0f;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int m = m0 + wm + 16 * i + 4 * lk + q;
        if (m < M) C[static_cast<size_t>(m) * ldc + n] = acc[i][j][q] + bn;
      }
    }
}

__global__ void gelu_fwd_kernel(const float* z, bf16* out, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = z[i];
  out[i] = f2bf(0.5f * x * (1.0f + erff(x * 0.70710678118654752f)));
}

__global__ void gelu_bwd_kernel(const float* z, float* dA, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = z[i];
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
  dA[i] *= cdf + x * pdf;
}

__global__ void to_bf16_kernel(const float* in, bf16* out, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = f2bf(in[i]);
}

__global__ void gather_kernel(const float* src, int stride, const int* idx, const float* mean,
                              const float* wz, int nin, int rows, bf16* x) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * nin) return;
  const int r = t / nin, i = t % nin;
  const float* row = src + static_cast<size_t>(r) * stride;
  x[t] = f2bf(whiten([&](int j) { return row[idx[j]]; }, mean, wz, nin, i));
}

// a training batch: rows order[b0..b0+nb) of the device-resident set
__global__ void gather_rows_kernel(const float* tx, const float* ty, const int* order, int b0,
                                   int nb, int nin, const float* mean, const float* wz,
                                   bf16* x, float* y) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nb * (nin + 1)) return;
  const int r = t / (nin + 1), c = t % (nin + 1);
  const int row = order[b0 + r];
  const float* xr = tx + static_cast<size_t>(row) * nin;
  if (c < nin)
    x[static_cast<size_t>(r) * nin + c] = f2bf(whiten([&](int j) { return xr[j]; }, mean, wz, nin, c));
  else
    y[r] = ty[row];
}

// column sums: a workgroup of 256 threads per 64 columns, 4 row phases
// reduced through LDS in a fixed order
__global__ __launch_bounds__(256) void colsum_kernel(const float* dz, int M, int N, float* g) {
  __shared__ float part[4][64];
  const int c = threadIdx.x & 63, ph = threadIdx.x >> 6;
  const int n = blockIdx.x * 64 + c;
  float s = 0.0f;
  if (n < N)
    for (int m = ph; m < M; m += 4) s += dz[static_cast<size_t>(m) * N + n];
  part[ph][c] = s;
  __syncthreads();
  if (ph == 0 && n < N) g[n] = (part[0][c] + part[1][c]) + (part[2][c] + part[3][c]);
}

__global__ void mse_grad_kernel(const float* pred, const float* y, int M, float* d, float* se) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const float r = pred[m] - y[m];
  d[m] = 2.0f * r / static_cast<float>(M);
  se[m] = r * r;
}

__global__ void adam_kernel(float* w, const float* g, float* m, float* v, bf16* wb, size_t n,
                            float lr, float b1, float b2, float eps, float c1, float c2) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  w[i] = wi;
  wb[i] = f2bf(wi);
}

__global__ void unstd_kernel(const float* pred, int n, double ym, double ys, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = static_cast<double>(pred[i]) * ys + ym;
}

__global__ __launch_bounds__(256) void pack_kernel(const bf16* wb, int H, int K, bf16* out) {
  const int Hp = (H + 15) & ~15, Kp = (K + 15) & ~15, nks = Kp / 16;
  const size_t n = static_cast<size_t>(Hp) * Kp;
  for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int q = static_cast<int>(e & 3), l = static_cast<int>((e >> 2) & 63);
    const size_t ts = e >> 8;
    const int sk = static_cast<int>(ts % nks), t = static_cast<int>(ts / nks);
    const int j = 16 * t + (l & 15), k = 16 * sk + 4 * (l >> 4) + q;
    out[e] = (j < H && k < K) ? wb[static_cast<size_t>(j) * K + k] : static_cast<bf16>(0);
  }
}

__global__ __launch_bounds__(kInfThreads) void infer_kernel(const float* src, int stride,
                                                             InferArgs a, int rows, double* out) {
  infer_rows(
      a, blockIdx.x * kInfRows, rows,
      [&](int r, int c) { return src[static_cast<size_t>(r) * stride + c]; }, out);
}

inline int ok() { return hipGetLastError() == hipSuccess ? 0 : 1; }
inline unsigned nb(size_t n, unsigned t) { return static_cast<unsigned>((n + t - 1) / t); }

}  // namespace

int gemm(int atype, const void* A, int lda, int ta, int btype, const void* B, int ldb, int tb,
         float* C, int ldc, const float* bias, int M, int N, int K, hipStream_t st) {
  const dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile);
  if (atype == 0 && btype == 0)
    hipLaunchKernelGGL((gemm_kernel<0, 0>), grid, dim3(256), 0, st, A, lda, ta, B, ldb, tb, C,
                       ldc, bias, M, N, K);
  else if (atype == 1 && btype == 0)
    hipLaunchKernelGGL((gemm_kernel<1, 0>), grid, dim3(256), 0, st, A, lda, ta, B, ldb, tb, C,
                       ldc, bias, M, N, K);
  else if (atype == 0 && btype == 1)
    hipLaunchKernelGGL((gemm_kernel<0, 1>), grid, dim3(256), 0, st, A, lda, ta, B, ldb, tb, C,
                       ldc, bias, M, N, K);
  else
    hipLaunchKernelGGL((gemm_kernel<1, 1>), grid, dim3(256), 0, st, A, lda, ta, B, ldb, tb, C,
                       ldc, bias, M, N, K);
  return ok();
}

int gelu_fwd(const float* z, bf16* out, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(gelu_fwd_kernel, dim3(nb(n, 256)), dim3(256), 0, st, z, out, n);
  return ok();
}

int gelu_bwd(const float* z, float* dA, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(gelu_bwd_kernel, dim3(nb(n, 256)), dim3(256), 0, st, z, dA, n);
  return ok();
}

int to_bf16(const float* in, bf16* out, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(to_bf16_kernel, dim3(nb(n, 256)), dim3(256), 0, st, in, out, n);
  return ok();
}

int gather_inputs(const float* src, int stride, const int* idx, const float* mean,
                  const float* wz, int nin, int rows, bf16* x, hipStream_t st) {
  const size_t n = static_cast<size_t>(rows) * nin;
  hipLaunchKernelGGL(gather_kernel, dim3(nb(n, 256)), dim3(256), 0, st, src, stride, idx, mean,
                     wz, nin, rows, x);
  return ok();
}

int gather_rows(const float* tx, const float* ty, const int* order, int b0, int nrows, int nin,
                const float* mean, const float* wz, bf16* x, float* y, hipStream_t st) {
  const size_t n = static_cast<size_t>(nrows) * (nin + 1);
  hipLaunchKernelGGL(gather_rows_kernel, dim3(nb(n, 256)), dim3(256), 0, st, tx, ty, order, b0,
                     nrows, nin, mean, wz, x, y);
  return ok();
}

int colsum(const float* dz, int M, int N, float* g, hipStream_t st) {
  hipLaunchKernelGGL(colsum_kernel, dim3(nb(N, 64)), dim3(256), 0, st, dz, M, N, g);
  return ok();
}

int mse_grad(const float* pred, const float* y, int M, float* d, float* se, hipStream_t st) {
  hipLaunchKernelGGL(mse_grad_kernel, dim3(nb(M, 256)), dim3(256), 0, st, pred, y, M, d, se);
  return ok();
}

int adam(float* w, const float* g, float* m, float* v, bf16* wb, size_t n, float lr, float b1,
         float b2, float eps, float c1, float c2, hipStream_t st) {
  hipLaunchKernelGGL(adam_kernel, dim3(nb(n, 256)), dim3(256), 0, st, w, g, m, v, wb, n, lr, b1,
                     b2, eps, c1, c2);
  return ok();
}

int pack_weights(const bf16* wb, int H, int K, bf16* out, hipStream_t st) {
  const size_t n = static_cast<size_t>((H + 15) & ~15) * ((K + 15) & ~15);
  const unsigned g = nb(n, 256);
  hipLaunchKernelGGL(pack_kernel, dim3(g < 1024 ? g : 1024), dim3(256), 0, st, wb, H, K, out);
  return ok();
}

int infer(const float* src, int stride, const InferArgs& a, int rows, double* out,
          hipStream_t st) {
  if (rows <= 0) return 0;
  hipLaunchKernelGGL(infer_kernel, dim3((rows + kInfRows - 1) / kInfRows), dim3(kInfThreads),
                     infer_lds_bytes(a.nin, a.H), st, src, stride, a, rows, out);
  return ok();
}

int unstandardize(const float* pred, int n, double ym, double ys, double* out, hipStream_t st) {
  hipLaunchKernelGGL(unstd_kernel, dim3(nb(n, 256)), dim3(256), 0, st, pred, n, ym, ys, out);
  return ok();
}

}  // namespace pofd_mlp_dev
