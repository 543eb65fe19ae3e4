// The surrogate network's fused inference as a device function, shared by
// the prediction kernel (mlp_kernel.hip) and the sampler's fused proposal
// and first-stage kernel (ensemble_kernel.hip).  Device code only.
#pragma once

#include <hip/hip_runtime.h>

#include "mlp_internal.h"

namespace pofd_mlp_dev {

__device__ __forceinline__ float bf2f(bf16 v) {
  return __uint_as_float(static_cast<uint32_t>(v) << 16);
}

__device__ __forceinline__ bf16 f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<bf16>((u >> 16) | 0x40);  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return static_cast<bf16>(u >> 16);
}

// whitened input c of one row: sum_j wz[c][j] (x_j - mean_j), float, j in
// order (a diagonal wz gives wz[c][c] (x_c - mean_c) exactly)
template <typename F>
__device__ __forceinline__ float whiten(F xj, const float* mean, const float* wz, int nin,
                                        int c) {
  float acc = 0.0f;
  for (int j = 0; j < nin; ++j) acc = fmaf(wz[c * nin + j], xj(j) - mean[j], acc);
  return acc;
}

// ---- fused inference: gather + standardise, three GELU layers on the
// matrix cores, output layer, de-standardisation, in one launch.  A
// workgroup of sixteen waves takes 16 rows; the activations stay in LDS as
// bfloat16 and the weights stream from L2.  Each wave owns 16-column tiles
// of the layer output (tiles w, w+16, ...) and accumulates them with
// v_mfma_f32_16x16x16_bf16: lane l holds A[l%16][4(l/16)..+3] and
// B[4(l/16)..+3][l%16] and receives D[4(l/16)..+3][l%16].  The weight
// fragments of sixteen k-steps are loaded before their products, so a wave
// waits on memory once per 256 inputs (the weights are usually evicted
// from L2 by the likelihood kernels between two predictions).
typedef short short4v __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));
constexpr int kInfRows = 16;
constexpr int kMaxIn = 64;  // network inputs (pofd_mlp_create's limit)
constexpr int kInfThreads = 1024;
constexpr int kInfWaves = kInfThreads / 64;

inline size_t infer_lds_bytes(int nin, int H) {
  const int Hp = (H + 15) & ~15, Kin = (nin + 15) & ~15;
  const int ld = (Hp > Kin ? Hp : Kin) + 4;
  return 2 * kInfRows * ld * sizeof(bf16);
}

// the network's predictions for rows r0 .. r0 + kInfRows - 1 (those below
// `rows`) of a matrix whose row r, column c is x(r, c), the inputs the
// columns a.idx[0 .. nin-1]: out[r] = net(whitened inputs) * ys + ym.  Run
// by all kInfThreads threads of a workgroup, with infer_lds_bytes() of
// dynamic LDS.
template <typename X>
__device__ void infer_rows(const InferArgs& a, const int r0, const int rows, X x, double* out) {
  extern __shared__ bf16 act[];  // two buffers of kInfRows x ld bf16
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nin = a.nin, H = a.H;
  const bf16 *wb = a.wb, *wpk = a.wpk;
  const float* w = a.w;
  const InferLayout& lay = a.lay;
  const int Hp = (H + 15) & ~15;
  const int Kin = (nin + 15) & ~15;
  const int ld = (Hp > Kin ? Hp : Kin) + 4;  // row pitch (bf16), 8-byte aligned
  bf16* bufA = act;
  bf16* bufB = act + kInfRows * ld;
  // inputs: the column map, the whitening and the rows' raw values staged
  // in LDS first (two dependent global reads in all), then whitened there
  __shared__ int sidx[kMaxIn];
  __shared__ float smean[kMaxIn];
  __shared__ float swz[kMaxIn * kMaxIn];
  __shared__ float sx[kInfRows * kMaxIn];
  for (int e = tid; e < nin * nin; e += kInfThreads) swz[e] = a.wz[e];
  if (tid < nin) {
    sidx[tid] = a.idx[tid];
    smean[tid] = a.mean[tid];
  }
  __syncthreads();
  for (int e = tid; e < kInfRows * nin; e += kInfThreads) {
    const int i = e / nin, j = e % nin;
    const int r = r0 + i;
    sx[e] = (r < rows) ? x(r, sidx[j]) : 0.0f;
  }
  __syncthreads();
  for (int e = tid; e < kInfRows * Kin; e += kInfThreads) {
    const int i = e / Kin, c = e % Kin;
    float v = 0.0f;
    if (c < nin && r0 + i < rows) v = whiten([&](int j) { return sx[i * nin + j]; }, smean, swz, nin, c);
    bufA[i * ld + c] = f2bf(v);
  }
  __syncthreads();
  const int li = lane & 15, lk = (lane >> 4) * 4;
  const int ntile = Hp / 16;
  for (int l = 0; l < 3; ++l) {
    const int Kp = (l == 0) ? Kin : Hp;
    const int nks = Kp / 16;
    const float* b = w + lay.b[l];
    for (int t = wave; t < ntile; t += kInfWaves) {
      const int j = t * 16 + li;  // output column this lane feeds B for
      float4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
      // this tile's fragments: 64 lanes x 8 contiguous bytes per k-step
      const short4v* wp =
          reinterpret_cast<const short4v*>(wpk + lay.p[l]) + static_cast<size_t>(t) * nks * 64 + lane;
      int s0 = 0;
      // sixteen k-steps at a time: their fragments all in flight, one wait
      for (; s0 + 16 <= nks; s0 += 16) {
        short4v bq[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) bq[u] = wp[(s0 + u) * 64];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const short4v a =
              *reinterpret_cast<const short4v*>(bufA + li * ld + 16 * (s0 + u) + lk);
          acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bq[u], acc, 0, 0, 0);
        }
      }
      for (; s0 < nks; ++s0) {
        const short4v bq = wp[s0 * 64];
        const short4v a = *reinterpret_cast<const short4v*>(bufA + li * ld + 16 * s0 + lk);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bq, acc, 0, 0, 0);
      }
      // D rows lk..lk+3, column j
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        float v = 0.0f;
        if (j < H) {
          const float z = acc[q] + b[j];
          v = 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
        }
        bufB[(lk + q) * ld + t * 16 + li] = f2bf(v);
      }
    }
    __syncthreads();
    bf16* tmp = bufA;
    bufA = bufB;
    bufB = tmp;
  }
  // output layer: 64 threads (one wave) per row
  {
    const int i = wave, s = lane;
    const bf16* W4 = wb + lay.w[3];
    float acc = 0.0f;
    for (int j = s; j < H; j += 64) acc += bf2f(bufA[i * ld + j]) * bf2f(W4[j]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o);
    const int r = r0 + i;
    if (s == 0 && r < rows) out[r] = static_cast<double>(acc + w[lay.b[3]]) * a.ys + a.ym;
  }
}


}  // namespace pofd_mlp_dev
