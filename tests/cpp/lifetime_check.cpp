// Object lifetimes of the C ABI, built with the library's sources into one
// executable whose host code carries AddressSanitizer and UBSan (device code
// unchanged): contexts whose workspaces grow, a captured multi-field graph
// made stale by that growth and created again, the allocation-free entry
// point, a two-band context whose workspace grows, and a sampler with a
// surrogate network (training, delayed acceptance, the surrogate transition)
// torn down in the documented order.  Any host memory error, leak aside,
// ends the program with the sanitizer's report; success prints
// "lifetime_check ok".
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "../../include/pofd_affine_amd.h"

#define CHECK(x)                                                                  \
  do {                                                                            \
    const int rc_ = (x);                                                          \
    if (rc_ != POFD_OK) {                                                         \
      std::fprintf(stderr, "%s:%d: %s returned %d\n", __FILE__, __LINE__, #x, rc_); \
      return 1;                                                                   \
    }                                                                             \
  } while (0)
// the same, with the context's message
#define CCHECK(c, x)                                                                   \
  do {                                                                                 \
    const int rc_ = (x);                                                               \
    if (rc_ != POFD_OK) {                                                              \
      std::fprintf(stderr, "%s:%d: %s returned %d: %s\n", __FILE__, __LINE__, #x, rc_, \
                   pofd_last_error(c));                                                \
      return 1;                                                                        \
    }                                                                                  \
  } while (0)
#define HCHECK(x)                                                                 \
  do {                                                                            \
    if ((x) != hipSuccess) {                                                      \
      std::fprintf(stderr, "%s:%d: %s failed\n", __FILE__, __LINE__, #x);         \
      return 1;                                                                   \
    }                                                                             \
  } while (0)

namespace {

const double kKnots[8] = {0.002, 0.005, 0.010, 0.020, 0.045, 0.100, 0.200, 0.450};

// broken power-law counts, dN/dS in Jy^-1 deg^-2
double dnds(double s) {
  const double n0 = 5e6, s0 = 0.010, sb = 0.020;
  if (s <= sb) return n0 * std::pow(s / s0, -2.2);
  return n0 * std::pow(sb / s0, -2.2) * std::pow(s / sb, -3.5);
}

// peak-normalised Gaussian beam of the given FWHM on a 6" grid of
// 2 half + 1 pixels a side (half from the FWHM when negative)
std::vector<double> gauss_beam(double fwhm, int half = -1) {
  if (half < 0) half = static_cast<int>(std::floor(4.5 * fwhm / 6.0));
  const double sig = fwhm / (2.0 * std::sqrt(2.0 * std::log(2.0))) / 6.0;
  std::vector<double> b;
  for (int y = -half; y <= half; ++y)
    for (int x = -half; x <= half; ++x) b.push_back(std::exp(-0.5 * (x * x + y * y) / (sig * sig)));
  return b;
}

// noise with a few bright pixels, mean subtracted
std::vector<double> noise_map(size_t n, double sigma, unsigned seed) {
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, sigma);
  std::vector<double> m(n);
  for (double& v : m) v = nd(rng);
  for (size_t i = 0; i < n; i += 97) m[i] += 0.02 * static_cast<double>(i % 5);
  double mean = 0.0;
  for (double v : m) mean += v;
  mean /= static_cast<double>(n);
  for (double& v : m) v -= mean;
  return m;
}

// [8 knot log10 values, sigma_mult, 2 bonus] rows near the generating model
std::vector<float> thetas1d(int nw, unsigned seed) {
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 0.05);
  std::vector<float> th(static_cast<size_t>(nw) * 11, 0.0f);
  for (int w = 0; w < nw; ++w) {
    for (int k = 0; k < 8; ++k) th[w * 11 + k] = static_cast<float>(std::log10(dnds(kKnots[k])) + nd(rng));
    th[w * 11 + 8] = 1.0f;
  }
  return th;
}

int make_ctx1d(pofd_ctx** ctx, uint32_t nflux, const std::vector<double>& map, bool binned) {
  CHECK(pofd_create(ctx, 0));
  CHECK(pofd1d_configure(*ctx, nflux, 256, kKnots, 8, 0.0));
  const std::vector<double> b = gauss_beam(18.1);
  pofd_beam* bm = nullptr;
  CHECK(pofd_beam_create(&bm, b.data(), b.size(), 6.0, 1e-6));
  CHECK(pofd_beam_make_histogram(bm, 120));
  int bs = -1, f = -1;
  CHECK(pofd1d_add_beamset(*ctx, bm, &bs));
  pofd_beam_destroy(bm);  // the context holds its own copy
  if (binned) {
    std::vector<uint32_t> bins(400);
    double c0 = 0.0, dl = 0.0;
    CHECK(pofd_data_bin(map.data(), map.size(), 400, bins.data(), &c0, &dl));
    CHECK(pofd1d_add_field(*ctx, bs, nullptr, map.size(), bins.data(), 400, c0, dl, 0.0025, 1.0, &f));
  } else {
    CHECK(pofd1d_add_field(*ctx, bs, map.data(), map.size(), nullptr, 0, 0.0, 0.0, 0.0025, 1.0, &f));
  }
  CHECK(pofd1d_prepare(*ctx));
  return 0;
}

int check_finite(const std::vector<double>& ll, const std::vector<int32_t>& st, const char* what) {
  for (size_t i = 0; i < ll.size(); ++i)
    if (st[i] != POFD_OK || !std::isfinite(ll[i])) {
      std::fprintf(stderr, "%s: walker %zu status %d logL %g\n", what, i, st[i], ll[i]);
      return 1;
    }
  return 0;
}

int one_band() {
  const std::vector<double> m1 = noise_map(150 * 150, 0.0025, 1), m2 = noise_map(90 * 90, 0.003, 2);
  pofd_ctx *a = nullptr, *b = nullptr;
  if (make_ctx1d(&a, 32768, m1, false) || make_ctx1d(&b, 16384, m2, true)) return 1;
  const int W = 16;
  std::vector<float> th = thetas1d(64, 3);
  std::vector<double> ll(64);
  std::vector<int32_t> st(64);
  CHECK(pofd1d_loglike_batch(a, th.data(), W, 11, ll.data(), st.data()));
  ll.resize(W);
  st.resize(W);
  if (check_finite(ll, st, "first batch")) return 1;

  // a graph over both fields
  float* d_th = nullptr;
  double* d_ll = nullptr;
  int32_t* d_st = nullptr;
  HCHECK(hipMalloc(&d_th, th.size() * sizeof(float)));
  HCHECK(hipMalloc(&d_ll, 64 * sizeof(double)));
  HCHECK(hipMalloc(&d_st, 64 * sizeof(int32_t)));
  HCHECK(hipMemcpy(d_th, th.data(), th.size() * sizeof(float), hipMemcpyHostToDevice));
  pofd_ctx* both[2] = {a, b};
  hipStream_t sm = nullptr;  // a capture needs a stream of its own
  HCHECK(hipStreamCreate(&sm));
  pofd_graph* g = nullptr;
  CCHECK(a, pofd1d_graph_create(both, 2, d_th, W, 11, d_ll, d_st, sm, &g));
  CCHECK(a, pofd1d_graph_launch(g, sm));
  CCHECK(a, pofd1d_graph_launch(g, sm));
  HCHECK(hipStreamSynchronize(sm));
  // a larger eager batch grows the workspace: the graph is stale
  std::vector<double> ll2(64);
  std::vector<int32_t> st2(64);
  CHECK(pofd1d_loglike_batch(a, th.data(), 64, 11, ll2.data(), st2.data()));
  if (pofd1d_graph_launch(g, sm) != POFD_E_ARG) {
    std::fprintf(stderr, "a stale graph launched\n");
    return 1;
  }
  std::printf("stale graph refused: %s\n", pofd_last_error(a));
  pofd1d_graph_destroy(g);
  g = nullptr;
  CCHECK(a, pofd1d_graph_create(both, 2, d_th, W, 11, d_ll, d_st, sm, &g));
  CCHECK(a, pofd1d_graph_launch(g, sm));
  HCHECK(hipStreamSynchronize(sm));
  pofd1d_graph_destroy(g);
  HCHECK(hipStreamDestroy(sm));

  // the allocation-free entry point and a forced split
  CHECK(pofd1d_loglike_batch_device_noalloc(a, d_th, 8, 11, d_ll, d_st, nullptr));
  HCHECK(hipDeviceSynchronize());
  CHECK(pofd1d_set_parts(a, 4));
  CHECK(pofd1d_loglike_batch(a, th.data(), 20, 11, ll2.data(), st2.data()));
  CHECK(pofd1d_set_parts(a, 0));
  const int32_t sizes[3] = {8, 40, 64};
  int32_t chosen[3] = {0, 0, 0};
  CHECK(pofd1d_autotune_parts(a, th.data(), 11, sizes, 3, 1, chosen));

  HCHECK(hipFree(d_th));
  HCHECK(hipFree(d_ll));
  HCHECK(hipFree(d_st));
  pofd_destroy(b);
  pofd_destroy(a);
  return 0;
}

int two_band() {
  // both beams on the grid of the wider one
  const int half = static_cast<int>(std::floor(4.5 * 25.2 / 6.0));
  const std::vector<double> b1 = gauss_beam(18.1, half), b2 = gauss_beam(25.2, half);
  pofd_dbeam* db = nullptr;
  CHECK(pofd_dbeam_create(&db, b1.data(), b2.data(), b1.size(), 6.0, 1e-6));
  CHECK(pofd_dbeam_make_histogram(db, 60));
  pofd_ctx* c = nullptr;
  CHECK(pofd_create(&c, 0));
  const double kp[6] = {0.003, 0.008, 0.020, 0.050, 0.120, 0.300};
  const double sp[3] = {0.005, 0.030, 0.200}, op[3] = {0.005, 0.030, 0.200};
  CHECK(pofd2d_configure(c, 256, 64, 1, kp, 6, sp, 3, op, 3, 0.0, 0.0, 0.0));
  int bs = -1, f = -1;
  CHECK(pofd2d_add_beamset(c, db, &bs));
  pofd_dbeam_destroy(db);
  const std::vector<double> d1 = noise_map(80 * 80, 0.0025, 5), d2 = noise_map(80 * 80, 0.0027, 6);
  CHECK(pofd2d_add_field(c, bs, d1.data(), d2.data(), d1.size(), nullptr, 0, 0, nullptr, 0.0025,
                         0.0027, 1.0, &f));
  const int npar = 6 + 3 + 3 + 2 + 4;
  std::vector<float> th(static_cast<size_t>(8) * npar, 0.0f);
  const float sv[3] = {0.1f, 0.2f, 0.3f}, ov[3] = {-0.2f, -0.05f, 0.1f};
  for (int w = 0; w < 8; ++w) {
    float* t = th.data() + w * npar;
    for (int k = 0; k < 6; ++k) t[k] = static_cast<float>(std::log10(dnds(kp[k])) + 0.01 * w);
    for (int k = 0; k < 3; ++k) t[6 + k] = sv[k];
    for (int k = 0; k < 3; ++k) t[9 + k] = ov[k];
    t[12] = t[13] = 1.0f;
  }
  CHECK(pofd2d_prepare(c, th.data(), npar));
  std::vector<double> ll(8);
  std::vector<int32_t> st(8);
  CHECK(pofd2d_loglike_batch(c, th.data(), 2, npar, ll.data(), st.data()));
  CHECK(pofd2d_loglike_batch(c, th.data(), 8, npar, ll.data(), st.data()));  // grows
  // (the per-walker status is the likelihood's answer for this small
  // synthetic set-up; what is checked here is the host code around it)
  std::printf("two-band statuses:");
  for (int w = 0; w < 8; ++w) std::printf(" %d", st[w]);
  std::printf("\n");
  pofd_destroy(c);
  return 0;
}

int sampler() {
  const std::vector<double> m = noise_map(120 * 120, 0.0025, 7);
  pofd_ctx* c = nullptr;
  if (make_ctx1d(&c, 16384, m, true)) return 1;
  const int W = 32;
  pofd_ens* e = nullptr;
  CHECK(pofd_ens_create(&e, c, POFD_TARGET_1D, W, 11, 2.0, 9));
  const uint8_t state[11] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 2, 2};
  CHECK(pofd_ens_set_param_state(e, state));
  float dexp = 0.0f;
  CHECK(pofd_ens_set_exponent_mode(e, 1, &dexp));
  std::vector<float> th = thetas1d(W, 11);
  CHECK(pofd_ens_set_positions(e, th.data()));
  CHECK(pofd_ens_run(e, 30, 1.0, 1));
  uint32_t n = 0;
  CHECK(pofd_ens_chain_length(e, &n));
  std::vector<float> ch(static_cast<size_t>(n) * W * 11);
  std::vector<double> cll(static_cast<size_t>(n) * W);
  CHECK(pofd_ens_get_chain(e, 0, n, ch.data(), cll.data()));
  std::vector<float> x(static_cast<size_t>(n) * W * 8);
  for (size_t r = 0; r < static_cast<size_t>(n) * W; ++r)
    for (int k = 0; k < 8; ++k) x[r * 8 + k] = ch[r * 11 + k];
  pofd_mlp* net = nullptr;
  CHECK(pofd_mlp_create(&net, 8, 64, 3));
  double tr = 0.0, ho = 0.0;
  CHECK(pofd_mlp_train(net, x.data(), cll.data(), static_cast<uint32_t>(n) * W, 5, 128, 1e-3f,
                       0.1f, 4, &tr, &ho));
  const uint32_t idx[8] = {0, 1, 2, 3, 4, 5, 6, 7};
  int on = 0;
  CHECK(pofd_ens_set_surrogate(e, net, idx, 8, 1e30, 0.0, &on));
  if (!on) {
    std::fprintf(stderr, "surrogate not attached\n");
    return 1;
  }
  CHECK(pofd_ens_run(e, 10, 1.0, 1));
  CHECK(pofd_ens_set_surrogate_steps(e, 3));
  CHECK(pofd_ens_run(e, 5, 1.0, 1));
  uint64_t np = 0, ne = 0;
  CHECK(pofd_ens_counts(e, &np, &ne));
  std::vector<float> cur(static_cast<size_t>(W) * 11);
  std::vector<double> ll(W);
  CHECK(pofd_ens_get_state(e, cur.data(), ll.data(), nullptr, nullptr));
  // the network outlives the sampler that uses it
  pofd_ens_destroy(e);
  pofd_mlp_destroy(net);
  pofd_destroy(c);
  return 0;
}

}  // namespace

int main() {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    std::fprintf(stderr, "no GPU\n");
    return 2;
  }
  if (one_band()) return 1;
  std::printf("one band ok\n");
  if (two_band()) return 1;
  std::printf("two bands ok\n");
  if (sampler()) return 1;
  std::printf("sampler ok\n");
  std::printf("lifetime_check ok\n");
  return 0;
}
