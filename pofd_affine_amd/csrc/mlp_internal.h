// Declarations shared by the surrogate network's host code and kernels.
#pragma once

#include <cstddef>
#include <cstdint>

#include <hip/hip_runtime.h>

namespace pofd_mlp_dev {

typedef uint16_t bf16;  // bfloat16 bit pattern

// C[m][n] = sum_k A(m,k) B(k,n) (+ bias[n]) with FP32 accumulation.
// A(m,k) = ta ? A[k*lda + m] : A[m*lda + k]; B(k,n) likewise with tb.
// Element types: bf16 (0) or float (1) for A and B; C is float.
int gemm(int atype, const void* A, int lda, int ta, int btype, const void* B, int ldb, int tb,
         float* C, int ldc, const float* bias, int M, int N, int K, hipStream_t st);

// out_bf16 = gelu(z) (z already includes the bias)
int gelu_fwd(const float* z, bf16* out, size_t n, hipStream_t st);
// dz = dA * gelu'(z), in place on dA
int gelu_bwd(const float* z, float* dA, size_t n, hipStream_t st);
// float -> bf16 (round to nearest even)
int to_bf16(const float* in, bf16* out, size_t n, hipStream_t st);
// whitened inputs: x[r][i] = sum_j wz[i][j] (src[r*stride + idx[j]] - mean[j]),
// as bf16 (wz row-major nin x nin; diagonal for plain standardisation)
int gather_inputs(const float* src, int stride, const int* idx, const float* mean,
                  const float* wz, int nin, int rows, bf16* x, hipStream_t st);
// training batch from a device-resident set: x[r] = whitened
// tx[order[b0 + r]] (bf16), y[r] = ty[order[b0 + r]]
int gather_rows(const float* tx, const float* ty, const int* order, int b0, int nb, int nin,
                const float* mean, const float* wz, bf16* x, float* y, hipStream_t st);
// column sums: g[n] = sum_m dz[m][n]
int colsum(const float* dz, int M, int N, float* g, hipStream_t st);
// output-layer loss gradient: d[m] = 2 (pred[m] - y[m]) / M, and the
// per-row squared errors into se[m]
int mse_grad(const float* pred, const float* y, int M, float* d, float* se, hipStream_t st);
// Adam step on master weights w with gradient g (Kingma and Ba, 2015), and
// the bf16 copy
int adam(float* w, const float* g, float* m, float* v, bf16* wb, size_t n, float lr, float b1,
         float b2, float eps, float c1, float c2, hipStream_t st);

// offsets (elements) of the weight and bias blocks in the parameter vector,
// and of the three hidden layers' weights in the packed copy (pack_weights)
struct InferLayout {
  size_t w[4], b[4], p[3];
};
// the packed copy of a hidden layer's bf16 weights W[j][k] (H outputs, K
// inputs, row-major at wb): the v_mfma_f32_16x16x16_bf16 B fragments in
// the order the inference kernel reads them, 16-column tile t, 16-input
// step s, lane l, element q at ((t * (Kp/16) + s) * 64 + l) * 4 + q =
// W[16 t + l % 16][16 s + 4 (l / 16) + q], zero past H and K (Hp x Kp
// values, Hp and Kp the 16-multiples at or above H and K)
int pack_weights(const bf16* wb, int H, int K, bf16* out, hipStream_t st);
// what the fused inference reads: the input columns idx, the whitening
// (mean, wz), the weights (hidden layers from the packed copy wpk, the
// output layer from wb, biases from w) and the target standardisation
struct InferArgs {
  const int* idx;
  const float* mean;
  const float* wz;
  int nin, H;
  const bf16* wb;
  const bf16* wpk;
  const float* w;
  InferLayout lay;
  double ym, ys;
};
// fused prediction of rows inputs gathered from src (stride floats per row,
// inputs at a.idx): out[r] = net(whitened x_r) * ys + ym
int infer(const float* src, int stride, const InferArgs& a, int rows, double* out,
          hipStream_t st);

// out[i] = pred[i] * ys + ym (double)
int unstandardize(const float* pred, int n, double ym, double ys, double* out, hipStream_t st);

}  // namespace pofd_mlp_dev

// the inference arguments of a network (mlp.cpp), its inputs the columns idx
// (device) of the rows it is given
struct pofd_mlp;
void pofd_mlp_infer_args(const pofd_mlp* s, const int* idx, pofd_mlp_dev::InferArgs* a);
