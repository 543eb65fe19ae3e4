// Host side of the log-likelihood surrogate (C ABI pofd_mlp_*): an MLP
// nin -> H -> H -> H -> 1 with GELU activations, bfloat16 weights and
// activations, float accumulation and float master weights, trained with
// Adam on standardised (theta, logL) pairs.  Used as the first stage of the
// delayed-acceptance stretch move (pofd_ens_set_surrogate).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../include/pofd_affine_amd.h"
#include "mlp_internal.h"

namespace md = pofd_mlp_dev;
using md::bf16;

struct pofd_mlp {
  int nin = 0, H = 0;
  size_t off[8] = {};   // W1 b1 W2 b2 W3 b3 W4 b4 in the parameter vector
  size_t npar = 0;
  float* w = nullptr;   // float master parameters
  bf16* wb = nullptr;   // bf16 copy (weights used by the GEMMs)
  bf16* wpk = nullptr;  // the hidden layers' bf16 weights packed for inference
  size_t pk_off[3] = {};
  float* g = nullptr;   // gradient
  float* m = nullptr;   // Adam moments
  float* v = nullptr;
  float* xmean = nullptr;  // input whitening (device): z = wz (x - xmean)
  float* xw = nullptr;     // nin x nin, row-major
  int* idx = nullptr;      // identity column map for host inputs
  double ymean = 0.0, ystd = 1.0;
  // batch scratch
  int cap = 0;
  bf16* X = nullptr;               // cap x nin
  float* Z[3] = {};                // cap x H pre-activations
  bf16* A[3] = {};                 // cap x H activations
  float* out = nullptr;            // cap
  float* y = nullptr;              // cap (standardised targets)
  float* dout = nullptr;           // cap
  float* se = nullptr;             // cap
  float* dA = nullptr;             // cap x H
  float* dA2 = nullptr;            // cap x H
  float* raw = nullptr;            // cap x nin raw inputs (host uploads)
  double* dpred = nullptr;         // cap predictions (log-likelihood units)
  hipStream_t stream = nullptr;
  long adam_t = 0;
  std::string err;
  ~pofd_mlp() {
    if (stream) (void)hipStreamSynchronize(stream);
    void* ps[] = {w,   wb,  wpk, g, m, v, xmean, xw, idx, X, Z[0], Z[1], Z[2], A[0], A[1],
                  A[2], out, y,   dout, se, dA, dA2, raw, dpred};
    for (void* p : ps)
      if (p) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

template <typename T>
int dalloc(T** p, size_t n) {
  return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(n, 1) * sizeof(T)) == hipSuccess
             ? 0
             : 1;
}

int ensure_cap(pofd_mlp* s, int n) {
  if (n <= s->cap) return 0;
  void* ps[] = {s->X, s->Z[0], s->Z[1], s->Z[2], s->A[0], s->A[1], s->A[2], s->out,
                s->y, s->dout, s->se,   s->dA,   s->dA2,  s->raw, s->dpred};
  (void)hipStreamSynchronize(s->stream);
  for (void* p : ps)
    if (p) (void)hipFree(p);
  const size_t N = static_cast<size_t>(n), H = s->H;
  int rc = dalloc(&s->X, N * s->nin) | dalloc(&s->out, N) | dalloc(&s->y, N) |
           dalloc(&s->dout, N) | dalloc(&s->se, N) | dalloc(&s->dA, N * H) |
           dalloc(&s->dA2, N * H) | dalloc(&s->raw, N * s->nin) | dalloc(&s->dpred, N);
  for (int l = 0; l < 3; ++l) rc |= dalloc(&s->Z[l], N * H) | dalloc(&s->A[l], N * H);
  if (rc) return 1;
  s->cap = n;
  return 0;
}

// forward pass of n rows from s->X into s->out (standardised prediction)
int forward(pofd_mlp* s, int n) {
  const int H = s->H;
  const bf16* in = s->X;
  int K = s->nin;
  int rc = 0;
  for (int l = 0; l < 3; ++l) {
    const bf16* W = s->wb + s->off[2 * l];
    const float* b = s->w + s->off[2 * l + 1];
    // Z = in . W^T + b: A(m,k) = in[m*K + k]; B(k,n) = W[n*K + k]
    rc |= md::gemm(0, in, K, 0, 0, W, K, 1, s->Z[l], H, b, n, H, K, s->stream);
    rc |= md::gelu_fwd(s->Z[l], s->A[l], static_cast<size_t>(n) * H, s->stream);
    in = s->A[l];
    K = H;
  }
  rc |= md::gemm(0, in, H, 0, 0, s->wb + s->off[6], H, 1, s->out, 1, s->w + s->off[7], n, 1, H,
                 s->stream);
  return rc;
}

// loss gradient of n rows (targets in s->y) into s->g
int backward(pofd_mlp* s, int n) {
  const int H = s->H;
  int rc = md::mse_grad(s->out, s->y, n, s->dout, s->se, s->stream);
  // output layer: gW4[0][j] = sum_m dout[m] A3[m][j]; gb4 = sum dout
  rc |= md::gemm(1, s->dout, n, 0, 0, s->A[2], H, 0, s->g + s->off[6], H, nullptr, 1, H, n,
                 s->stream);
  rc |= md::colsum(s->dout, n, 1, s->g + s->off[7], s->stream);
  // dA3 = dout . W4
  rc |= md::gemm(1, s->dout, 1, 0, 0, s->wb + s->off[6], H, 0, s->dA, H, nullptr, n, H, 1,
                 s->stream);
  float* dA = s->dA;
  float* dnext = s->dA2;
  for (int l = 2; l >= 0; --l) {
    rc |= md::gelu_bwd(s->Z[l], dA, static_cast<size_t>(n) * H, s->stream);  // dZ
    const int K = (l == 0) ? s->nin : H;
    const bf16* in = (l == 0) ? s->X : s->A[l - 1];
    // gW[o][i] = sum_m dZ[m][o] in[m][i]
    rc |= md::gemm(1, dA, H, 1, 0, in, K, 0, s->g + s->off[2 * l], K, nullptr, H, K, n,
                   s->stream);
    rc |= md::colsum(dA, n, H, s->g + s->off[2 * l + 1], s->stream);
    if (l > 0) {
      // dA_prev[m][i] = sum_o dZ[m][o] W[o][i]
      rc |= md::gemm(1, dA, H, 0, 0, s->wb + s->off[2 * l], H, 0, dnext, H, nullptr, n, H, H,
                     s->stream);
      std::swap(dA, dnext);
    }
  }
  return rc;
}

int upload_inputs(pofd_mlp* s, const float* x, int n) {
  if (ensure_cap(s, n)) return 1;
  if (hipMemcpyAsync(s->raw, x, static_cast<size_t>(n) * s->nin * sizeof(float),
                     hipMemcpyHostToDevice, s->stream) != hipSuccess)
    return 1;
  return md::gather_inputs(s->raw, s->nin, s->idx, s->xmean, s->xw, s->nin, n, s->X,
                           s->stream);
}

int upload_targets(pofd_mlp* s, const double* y, int n) {
  std::vector<float> ys(n);
  for (int i = 0; i < n; ++i) ys[i] = static_cast<float>((y[i] - s->ymean) / s->ystd);
  if (hipMemcpyAsync(s->y, ys.data(), n * sizeof(float), hipMemcpyHostToDevice, s->stream) !=
      hipSuccess)
    return 1;
  return hipStreamSynchronize(s->stream) == hipSuccess ? 0 : 1;
}

int mfail(pofd_mlp* s, int code, const std::string& msg) {
  if (s) s->err = msg;
  return code;
}

// the packed copy of the hidden layers' weights, after every change of wb
int repack(pofd_mlp* s) {
  int rc = 0;
  for (int l = 0; l < 3; ++l)
    rc |= md::pack_weights(s->wb + s->off[2 * l], s->H, l == 0 ? s->nin : s->H,
                           s->wpk + s->pk_off[l], s->stream);
  return rc;
}

}  // namespace

// used by the ensemble sampler: predictions (log likelihood units) of the
// rows of src (stride floats per row, inputs at idx) into out, on stream st
void pofd_mlp_infer_args(const pofd_mlp* s, const int* idx, md::InferArgs* a) {
  a->idx = idx;
  a->mean = s->xmean;
  a->wz = s->xw;
  a->nin = s->nin;
  a->H = s->H;
  a->wb = s->wb;
  a->wpk = s->wpk;
  a->w = s->w;
  for (int l = 0; l < 4; ++l) {
    a->lay.w[l] = s->off[2 * l];
    a->lay.b[l] = s->off[2 * l + 1];
  }
  for (int l = 0; l < 3; ++l) a->lay.p[l] = s->pk_off[l];
  a->ym = s->ymean;
  a->ys = s->ystd;
}

// used by the ensemble sampler: predictions (log likelihood units) of the
// rows of src (stride floats per row, inputs at idx) into out, on stream st
int pofd_mlp_predict_device(pofd_mlp* s, const float* src, int stride, const int* idx, int rows,
                            double* out, hipStream_t st) {
  md::InferArgs a;
  pofd_mlp_infer_args(s, idx, &a);
  return md::infer(src, stride, a, rows, out, st);
}

extern "C" {

int pofd_mlp_create(pofd_mlp** out, uint32_t nin, uint32_t hidden, uint64_t seed) {
  if (!out) return POFD_E_ARG;
  *out = nullptr;
  // the fused inference kernel keeps two 16-row activation tiles in LDS
  if (nin == 0 || nin > 64 || hidden == 0 || hidden > 512) return POFD_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return POFD_E_NOGPU;
  pofd_mlp* s = new pofd_mlp();
  s->nin = static_cast<int>(nin);
  s->H = static_cast<int>(hidden);
  const size_t H = hidden, I = nin;
  const size_t sz[8] = {H * I, H, H * H, H, H * H, H, H, 1};
  size_t o = 0;
  for (int k = 0; k < 8; ++k) {
    s->off[k] = o;
    o += (sz[k] + 7) & ~size_t(7);
  }
  s->npar = o;
  const size_t Hp = (H + 15) & ~size_t(15), Ip = (I + 15) & ~size_t(15);
  s->pk_off[0] = 0;
  s->pk_off[1] = Hp * Ip;
  s->pk_off[2] = Hp * Ip + Hp * Hp;
  if (hipStreamCreate(&s->stream) != hipSuccess ||
      (dalloc(&s->w, o) | dalloc(&s->wb, o) | dalloc(&s->wpk, Hp * Ip + 2 * Hp * Hp) |
       dalloc(&s->g, o) | dalloc(&s->m, o) |
       dalloc(&s->v, o) | dalloc(&s->xmean, I) | dalloc(&s->xw, I * I) | dalloc(&s->idx, I))) {
    delete s;
    return POFD_E_INTERNAL;
  }
  // He initialisation of the weights, zero biases
  std::mt19937_64 rng(seed);
  std::vector<float> w(o, 0.0f);
  for (int l = 0; l < 4; ++l) {
    const size_t fan_in = (l == 0) ? I : H;
    std::normal_distribution<double> nd(0.0, std::sqrt((l == 3 ? 1.0 : 2.0) / fan_in));
    const size_t n = (l == 0) ? H * I : (l == 3 ? H : H * H);
    for (size_t i = 0; i < n; ++i) w[s->off[2 * l] + i] = static_cast<float>(nd(rng));
  }
  std::vector<float> zeros(o, 0.0f), mean(I, 0.0f), eye(I * I, 0.0f);
  for (size_t i = 0; i < I; ++i) eye[i * I + i] = 1.0f;
  std::vector<int> idx(I);
  std::iota(idx.begin(), idx.end(), 0);
  hipError_t he = hipMemcpy(s->w, w.data(), o * sizeof(float), hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemcpy(s->m, zeros.data(), o * sizeof(float), hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemcpy(s->v, zeros.data(), o * sizeof(float), hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemcpy(s->g, zeros.data(), o * sizeof(float), hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemcpy(s->xmean, mean.data(), I * sizeof(float), hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemcpy(s->xw, eye.data(), I * I * sizeof(float), hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemcpy(s->idx, idx.data(), I * sizeof(int), hipMemcpyHostToDevice);
  if (he != hipSuccess || md::to_bf16(s->w, s->wb, o, s->stream) || repack(s) ||
      hipStreamSynchronize(s->stream) != hipSuccess) {
    delete s;
    return POFD_E_INTERNAL;
  }
  *out = s;
  return POFD_OK;
}

void pofd_mlp_destroy(pofd_mlp* s) { delete s; }

const char* pofd_mlp_last_error(const pofd_mlp* s) { return s ? s->err.c_str() : ""; }

int pofd_mlp_info(const pofd_mlp* s, uint32_t* nin, uint32_t* hidden) {
  if (!s) return POFD_E_ARG;
  if (nin) *nin = static_cast<uint32_t>(s->nin);
  if (hidden) *hidden = static_cast<uint32_t>(s->H);
  return POFD_OK;
}

int pofd_mlp_nparams(const pofd_mlp* s, uint32_t* n) {
  if (!s || !n) return POFD_E_ARG;
  *n = static_cast<uint32_t>(s->npar);
  return POFD_OK;
}

int pofd_mlp_layout(const pofd_mlp* s, uint64_t* offsets) {
  if (!s || !offsets) return POFD_E_ARG;
  for (int k = 0; k < 8; ++k) offsets[k] = s->off[k];
  return POFD_OK;
}

int pofd_mlp_get_params(pofd_mlp* s, float* w) {
  if (!s || !w) return POFD_E_ARG;
  return hipMemcpy(w, s->w, s->npar * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess
             ? POFD_OK
             : POFD_E_INTERNAL;
}

int pofd_mlp_set_params(pofd_mlp* s, const float* w) {
  if (!s || !w) return POFD_E_ARG;
  if (hipMemcpy(s->w, w, s->npar * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      md::to_bf16(s->w, s->wb, s->npar, s->stream) || repack(s) ||
      hipStreamSynchronize(s->stream) != hipSuccess)
    return POFD_E_INTERNAL;
  return POFD_OK;
}

int pofd_mlp_set_whitening(pofd_mlp* s, const float* xmean, const float* wmat, double ymean,
                           double ystd) {
  if (!s || !xmean || !wmat || !(ystd > 0.0)) return POFD_E_ARG;
  const size_t I = s->nin;
  for (size_t i = 0; i < I * I; ++i)
    if (!std::isfinite(wmat[i]))
      return mfail(s, POFD_E_ARG, "pofd_mlp_set_whitening: non-finite matrix");
  if (hipStreamSynchronize(s->stream) != hipSuccess ||
      hipMemcpy(s->xmean, xmean, I * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(s->xw, wmat, I * I * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return POFD_E_INTERNAL;
  s->ymean = ymean;
  s->ystd = ystd;
  return POFD_OK;
}

int pofd_mlp_set_standardization(pofd_mlp* s, const float* xmean, const float* xstd,
                                 double ymean, double ystd) {
  if (!s || !xmean || !xstd || !(ystd > 0.0)) return POFD_E_ARG;
  std::vector<float> wz(static_cast<size_t>(s->nin) * s->nin, 0.0f);
  for (int i = 0; i < s->nin; ++i) {
    if (!(xstd[i] > 0.0f)) return mfail(s, POFD_E_ARG, "pofd_mlp_set_standardization: std <= 0");
    wz[static_cast<size_t>(i) * s->nin + i] = 1.0f / xstd[i];
  }
  return pofd_mlp_set_whitening(s, xmean, wz.data(), ymean, ystd);
}

int pofd_mlp_predict(pofd_mlp* s, const float* x, uint32_t n, double* out) {
  if (!s || !x || !out) return POFD_E_ARG;
  if (n == 0) return POFD_OK;
  if (ensure_cap(s, static_cast<int>(n)) ||
      hipMemcpyAsync(s->raw, x, static_cast<size_t>(n) * s->nin * sizeof(float),
                     hipMemcpyHostToDevice, s->stream) != hipSuccess ||
      pofd_mlp_predict_device(s, s->raw, s->nin, s->idx, static_cast<int>(n), s->dpred,
                              s->stream) ||
      hipMemcpyAsync(out, s->dpred, n * sizeof(double), hipMemcpyDeviceToHost, s->stream) !=
          hipSuccess ||
      hipStreamSynchronize(s->stream) != hipSuccess)
    return mfail(s, POFD_E_INTERNAL, "pofd_mlp_predict: device failure");
  return POFD_OK;
}

int pofd_mlp_loss_grad(pofd_mlp* s, const float* x, const double* y, uint32_t n, float* grad,
                       double* loss) {
  if (!s || !x || !y || n == 0) return POFD_E_ARG;
  if (upload_inputs(s, x, static_cast<int>(n)) || upload_targets(s, y, static_cast<int>(n)) ||
      forward(s, static_cast<int>(n)) || backward(s, static_cast<int>(n)))
    return mfail(s, POFD_E_INTERNAL, "pofd_mlp_loss_grad: device failure");
  std::vector<float> se(n);
  if (hipMemcpyAsync(se.data(), s->se, n * sizeof(float), hipMemcpyDeviceToHost, s->stream) !=
          hipSuccess ||
      (grad && hipMemcpyAsync(grad, s->g, s->npar * sizeof(float), hipMemcpyDeviceToHost,
                              s->stream) != hipSuccess) ||
      hipStreamSynchronize(s->stream) != hipSuccess)
    return POFD_E_INTERNAL;
  if (loss) {
    double t = 0.0;
    for (float e : se) t += e;
    *loss = t / n;
  }
  return POFD_OK;
}

int pofd_mlp_train(pofd_mlp* s, const float* x, const double* y, uint32_t n, uint32_t epochs,
                   uint32_t batch, float lr, float holdout_frac, uint64_t seed,
                   double* train_rms, double* holdout_rms) {
  if (!s || !x || !y || n < 4 || batch == 0 || !(lr > 0.0f) || holdout_frac < 0.0f ||
      holdout_frac >= 1.0f)
    return POFD_E_ARG;
  const int nin = s->nin;
  // shuffle once; the last holdout_frac of the rows are held out
  std::mt19937_64 rng(seed);
  std::vector<uint32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0u);
  std::shuffle(perm.begin(), perm.end(), rng);
  const uint32_t nho = static_cast<uint32_t>(holdout_frac * n);
  const uint32_t ntr = n - nho;
  // whitening from the training rows: z = L^-1 (x - mean) with L L^T the
  // sample covariance (Cholesky), so that the inputs are uncorrelated with
  // unit variance before their bfloat16 rounding -- strongly correlated
  // parameters would otherwise reach the network as small differences of
  // rounded numbers
  std::vector<double> mu(nin, 0.0), cov(static_cast<size_t>(nin) * nin, 0.0);
  double ym = 0.0, yv = 0.0;
  for (uint32_t r = 0; r < ntr; ++r) {
    const uint32_t i = perm[r];
    for (int c = 0; c < nin; ++c) mu[c] += x[static_cast<size_t>(i) * nin + c];
    ym += y[i];
  }
  for (int c = 0; c < nin; ++c) mu[c] /= ntr;
  ym /= ntr;
  for (uint32_t r = 0; r < ntr; ++r) {
    const uint32_t i = perm[r];
    const float* xr = x + static_cast<size_t>(i) * nin;
    for (int a = 0; a < nin; ++a)
      for (int b = 0; b <= a; ++b) cov[a * nin + b] += (xr[a] - mu[a]) * (xr[b] - mu[b]);
    yv += (y[i] - ym) * (y[i] - ym);
  }
  std::vector<float> xm(nin), wz(static_cast<size_t>(nin) * nin, 0.0f);
  for (int c = 0; c < nin; ++c) xm[c] = static_cast<float>(mu[c]);
  {
    double tr = 0.0;
    for (int a = 0; a < nin; ++a) tr += cov[a * nin + a] / ntr;
    const double jitter = 1e-12 * (tr > 0 ? tr / nin : 1.0);
    std::vector<double> L(static_cast<size_t>(nin) * nin, 0.0);
    bool ok = true;
    for (int a = 0; a < nin && ok; ++a) {
      for (int b = 0; b <= a; ++b) {
        double v = cov[a * nin + b] / ntr + (a == b ? jitter : 0.0);
        for (int k = 0; k < b; ++k) v -= L[a * nin + k] * L[b * nin + k];
        if (a == b) {
          if (!(v > 0.0)) {
            ok = false;
            break;
          }
          L[a * nin + a] = std::sqrt(v);
        } else {
          L[a * nin + b] = v / L[b * nin + b];
        }
      }
    }
    if (ok) {
      // wz = L^-1 by forward substitution, column by column
      std::vector<double> Li(static_cast<size_t>(nin) * nin, 0.0);
      for (int c = 0; c < nin; ++c)
        for (int a = c; a < nin; ++a) {
          double v = (a == c) ? 1.0 : 0.0;
          for (int k = c; k < a; ++k) v -= L[a * nin + k] * Li[k * nin + c];
          Li[a * nin + c] = v / L[a * nin + a];
        }
      for (size_t i = 0; i < Li.size(); ++i) wz[i] = static_cast<float>(Li[i]);
    } else {  // a degenerate input: per-input standardisation
      for (int c = 0; c < nin; ++c) {
        const double sd = std::sqrt(cov[c * nin + c] / ntr);
        wz[c * nin + c] = static_cast<float>(sd > 0 ? 1.0 / sd : 1.0);
      }
    }
  }
  const double ysd = std::sqrt(yv / ntr);
  int rc = pofd_mlp_set_whitening(s, xm.data(), wz.data(), ym, ysd > 0 ? ysd : 1.0);
  if (rc) return rc;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  // the training rows stay on the device; each epoch uploads one shuffled
  // order and the batches are gathered there
  std::vector<float> hx(static_cast<size_t>(ntr) * nin), hy(ntr);
  for (uint32_t r = 0; r < ntr; ++r) {
    const uint32_t i = perm[r];
    std::memcpy(&hx[static_cast<size_t>(r) * nin], &x[static_cast<size_t>(i) * nin],
                nin * sizeof(float));
    hy[r] = static_cast<float>((y[i] - s->ymean) / s->ystd);
  }
  float* tx = nullptr;
  float* ty = nullptr;
  int* dorder = nullptr;
  auto release = [&]() {
    (void)hipStreamSynchronize(s->stream);
    if (tx) (void)hipFree(tx);
    if (ty) (void)hipFree(ty);
    if (dorder) (void)hipFree(dorder);
  };
  std::vector<int> order(ntr);
  std::iota(order.begin(), order.end(), 0);
  if (ensure_cap(s, static_cast<int>(std::min(batch, ntr))) || dalloc(&tx, hx.size()) ||
      dalloc(&ty, ntr) || dalloc(&dorder, static_cast<size_t>(ntr) * epochs) ||
      hipMemcpyAsync(tx, hx.data(), hx.size() * sizeof(float), hipMemcpyHostToDevice,
                     s->stream) != hipSuccess ||
      hipMemcpyAsync(ty, hy.data(), ntr * sizeof(float), hipMemcpyHostToDevice, s->stream) !=
          hipSuccess) {
    release();
    return mfail(s, POFD_E_INTERNAL, "pofd_mlp_train: device allocation failed");
  }
  // all epochs' orders up front (one copy, no reuse of a host buffer while
  // a copy may still read it)
  std::vector<int> orders(static_cast<size_t>(ntr) * epochs);
  for (uint32_t ep = 0; ep < epochs; ++ep) {
    std::shuffle(order.begin(), order.end(), rng);
    std::copy(order.begin(), order.end(), orders.begin() + static_cast<size_t>(ep) * ntr);
  }
  if (epochs > 0 &&
      hipMemcpyAsync(dorder, orders.data(), orders.size() * sizeof(int), hipMemcpyHostToDevice,
                     s->stream) != hipSuccess) {
    release();
    return mfail(s, POFD_E_INTERNAL, "pofd_mlp_train: order upload failed");
  }
  // cosine decay of the step size to 5% of lr over the run
  const uint32_t nbatch = (ntr + batch - 1) / batch;
  const double total = static_cast<double>(nbatch) * epochs;
  uint64_t it = 0;
  for (uint32_t ep = 0; ep < epochs; ++ep) {
    for (uint32_t b0 = 0; b0 < ntr; b0 += batch, ++it) {
      const uint32_t nb = std::min(batch, ntr - b0);
      if (md::gather_rows(tx, ty, dorder + static_cast<size_t>(ep) * ntr, static_cast<int>(b0),
                          static_cast<int>(nb), nin, s->xmean, s->xw, s->X, s->y,
                          s->stream) ||
          forward(s, nb) || backward(s, nb)) {
        release();
        return mfail(s, POFD_E_INTERNAL, "pofd_mlp_train: device failure");
      }
      ++s->adam_t;
      const float c1 = 1.0f / (1.0f - std::pow(b1, static_cast<float>(s->adam_t)));
      const float c2 = 1.0f / (1.0f - std::pow(b2, static_cast<float>(s->adam_t)));
      const double frac = static_cast<double>(it) / total;
      const float lrt =
          static_cast<float>(lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(M_PI * frac))));
      if (md::adam(s->w, s->g, s->m, s->v, s->wb, s->npar, lrt, b1, b2, eps, c1, c2,
                   s->stream)) {
        release();
        return mfail(s, POFD_E_INTERNAL, "pofd_mlp_train: adam launch failed");
      }
    }
  }
  release();
  if (repack(s)) return mfail(s, POFD_E_INTERNAL, "pofd_mlp_train: packing failed");
  // errors in log-likelihood units
  auto rms = [&](uint32_t first, uint32_t count, double* res) -> int {
    if (!res) return 0;
    if (count == 0) {
      *res = NAN;
      return 0;
    }
    std::vector<float> xx(static_cast<size_t>(count) * nin);
    std::vector<double> pred(count);
    for (uint32_t r = 0; r < count; ++r)
      std::memcpy(&xx[static_cast<size_t>(r) * nin],
                  &x[static_cast<size_t>(perm[first + r]) * nin], nin * sizeof(float));
    if (pofd_mlp_predict(s, xx.data(), count, pred.data())) return 1;
    double t = 0.0;
    for (uint32_t r = 0; r < count; ++r) {
      const double d = pred[r] - y[perm[first + r]];
      t += d * d;
    }
    *res = std::sqrt(t / count);
    return 0;
  };
  if (rms(0, ntr, train_rms) || rms(ntr, nho, holdout_rms))
    return mfail(s, POFD_E_INTERNAL, "pofd_mlp_train: evaluation failed");
  return POFD_OK;
}

}  // extern "C"
