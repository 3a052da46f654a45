// net.cpp -- the learned warm start: network object, training loop, and the
// C ABI entry points cdiff_net_* (include/cdiff.h).
#include "net.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>

#include "solver.h"

using namespace cdiff;

#define HIP_TRY(expr)                                 \
  do {                                                \
    hipError_t _e = (expr);                           \
    if (_e != hipSuccess) return hip_fail(_e, #expr); \
  } while (0)

#define LAUNCH_CHECK(name)                                        \
  do {                                                            \
    hipError_t _e = hipGetLastError();                            \
    if (_e != hipSuccess) return hip_fail(_e, "launch of " name); \
  } while (0)

int cdiff_net::fail(int code, const std::string& msg) {
  err = msg;
  return code;
}

int cdiff_net::hip_fail(hipError_t e, const char* what) {
  err = std::string(what) + ": " + hipGetErrorString(e);
  return CDIFF_HIP_ERROR;
}

int cdiff_net::init(uint64_t seed) {
  HIP_TRY(hipStreamCreate(&stream));
  auto alloc = [&](void** p, size_t bytes) -> hipError_t {
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) {
      owned.push_back(*p);
      e = hipMemsetAsync(*p, 0, bytes, stream);
    }
    return e;
  };
  HIP_TRY(alloc((void**)&params, sizeof(float) * NP_TOTAL));
  HIP_TRY(alloc((void**)&grad, sizeof(float) * NP_TOTAL));
  HIP_TRY(alloc((void**)&m1, sizeof(float) * NP_TOTAL));
  HIP_TRY(alloc((void**)&m2, sizeof(float) * NP_TOTAL));
  for (int l = 1; l <= NET_CONVS; ++l) {
    HIP_TRY(alloc((void**)&w.wp[l], sizeof(__bf16) * NET_C * 9 * NET_C));
    if (l >= 2) HIP_TRY(alloc((void**)&w.wf[l], sizeof(__bf16) * NET_C * 9 * NET_C));
  }
  HIP_TRY(alloc((void**)&w.wh, sizeof(__bf16) * NET_HO * 9 * NET_C));
  w.params = params;
  // He initialisation of the ReLU layers (the first layer's fan-in counts the
  // 24 features of a block that carry data), a small head: the first
  // predictions are near zero (the warm start then falls back to T^n)
  std::vector<float> p(NP_TOTAL, 0.f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.f, 1.f);
  auto fill = [&](int off, int n, float sd) {
    for (int i = 0; i < n; ++i) p[off + i] = sd * nd(rng);
  };
  fill(np_w(1), NET_C * 9 * NET_C, std::sqrt(2.f / (9 * 24)));
  for (int l = 2; l <= NET_CONVS; ++l) fill(np_w(l), NET_C * 9 * NET_C, std::sqrt(2.f / (9 * NET_C)));
  fill(NP_WH, NET_HO * 9 * NET_C, 0.01f * std::sqrt(1.f / (9 * NET_C)));
  HIP_TRY(hipMemcpyAsync(params, p.data(), sizeof(float) * NP_TOTAL, hipMemcpyHostToDevice, stream));
  launch_pack_weights(params, w, stream);
  LAUNCH_CHECK("k_pack_weights");
  HIP_TRY(hipStreamSynchronize(stream));
  return CDIFF_OK;
}

int cdiff_net::workspace(Img m) {
  if (m.H == wimg.H && m.W == wimg.W && XS) return CDIFF_OK;
  if (m.H < 1 || m.W < 1) return fail(CDIFF_BAD_ARG, "image size must be positive");
  HIP_TRY(hipStreamSynchronize(stream));
  for (void* p : ws) (void)hipFree(p);
  ws.clear();
  const int64_t npx = img_pixels(m), ni = (int64_t)m.H * m.W;
  auto alloc = [&](void** p, size_t bytes) -> hipError_t {
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) {
      ws.push_back(*p);
      e = hipMemsetAsync(*p, 0, bytes, stream);  // zero borders stay zero
    }
    return e;
  };
  const Img lv1 = img_half(m), lv2 = img_half(lv1), lv3 = img_half(lv2);
  const int64_t n1 = img_pixels(lv1), n2 = img_pixels(lv2), n3 = img_pixels(lv3);
  HIP_TRY(alloc((void**)&X0, sizeof(__bf16) * npx * NET_C0));
  for (__bf16** b : {&XS, &E1, &A6, &U1, &GU1, &GA6, &GE1}) HIP_TRY(alloc((void**)b, sizeof(__bf16) * n1 * NET_C));
  for (__bf16** b : {&P1, &E2, &A5, &U2, &GU2, &GA5, &GE2, &GP1})
    HIP_TRY(alloc((void**)b, sizeof(__bf16) * n2 * NET_C));
  for (__bf16** b : {&P2, &E3, &B3, &GB3, &GE3, &GP2}) HIP_TRY(alloc((void**)b, sizeof(__bf16) * n3 * NET_C));
  HIP_TRY(alloc((void**)&delta, sizeof(float) * ni));
  HIP_TRY(alloc((void**)&tgt, sizeof(float) * ni));
  HIP_TRY(alloc((void**)&G4, sizeof(float) * n1 * NET_HO));
  HIP_TRY(alloc((void**)&loss_part, sizeof(double) * 2 * loss_parts(m)));
  HIP_TRY(alloc((void**)&loss_sums, sizeof(double) * 2));
  HIP_TRY(alloc((void**)&res, sizeof(float) * ni));
  // weight-gradient partials: the largest of the conv layers (level 1) and the head
  const int64_t conv_part = (int64_t)conv_wgrad_blocks(lv1) * NET_C * conv_wgrad_npad();
  const int64_t head_part = (int64_t)head_wgrad_blocks(lv1) * NET_HO * (9 * NET_C + 1);
  HIP_TRY(alloc((void**)&part, sizeof(float) * std::max(conv_part, head_part)));
  HIP_TRY(hipStreamSynchronize(stream));
  wimg = m;
  return CDIFF_OK;
}

// the encoder-decoder up to U1 (learn.h)
static void encode_decode(cdiff_net& n, const __bf16* XSin, hipStream_t s) {
  const Img lv1 = img_half(n.wimg), lv2 = img_half(lv1), lv3 = img_half(lv2);
  const NetWeights& w = n.w;
  const float* P = n.params;
  launch_conv3x3(0, lv1, XSin, w.wp[1], P + np_b(1), nullptr, n.E1, s);
  launch_pool(lv1, n.E1, n.P1, s);
  launch_conv3x3(0, lv2, n.P1, w.wp[2], P + np_b(2), nullptr, n.E2, s);
  launch_pool(lv2, n.E2, n.P2, s);
  launch_conv3x3(0, lv3, n.P2, w.wp[3], P + np_b(3), nullptr, n.E3, s);
  launch_conv3x3(0, lv3, n.E3, w.wp[4], P + np_b(4), nullptr, n.B3, s);
  launch_up_add(lv2, n.B3, n.E2, n.A5, s);
  launch_conv3x3(0, lv2, n.A5, w.wp[5], P + np_b(5), nullptr, n.U2, s);
  launch_up_add(lv1, n.U2, n.E1, n.A6, s);
  launch_conv3x3(0, lv1, n.A6, w.wp[6], P + np_b(6), nullptr, n.U1, s);
}

int cdiff_net::forward(const __bf16* XSin, hipStream_t s) {
  encode_decode(*this, XSin, s);
  launch_head_infer(wimg, U1, w.wh, params, XSin, gate, delta, s);
  LAUNCH_CHECK("network forward");
  return CDIFF_OK;
}

int cdiff_net::backward(const __bf16* XSin, const float* target, double* loss_dev, hipStream_t s) {
  const Img lv0 = wimg, lv1 = img_half(lv0), lv2 = img_half(lv1), lv3 = img_half(lv2);
  encode_decode(*this, XSin, s);
  launch_head_fwd(lv0, U1, w.wh, params, XSin, gate, delta, s);
  launch_loss(loss_kind, lv0, XSin, delta, target, gate, res, loss_part, loss_sums, loss_dev, G4, s);
  // backward data (the ReLU derivatives applied where the skips join)
  launch_head_bwd(lv0, G4, w.wh, U1, GU1, s);
  launch_conv3x3(2, lv1, GU1, w.wf[6], nullptr, nullptr, GA6, s);
  launch_grad_up(lv1, GA6, U2, GU2, s);
  launch_conv3x3(2, lv2, GU2, w.wf[5], nullptr, nullptr, GA5, s);
  launch_grad_up(lv2, GA5, B3, GB3, s);
  launch_conv3x3(1, lv3, GB3, w.wf[4], nullptr, E3, GE3, s);
  launch_conv3x3(2, lv3, GE3, w.wf[3], nullptr, nullptr, GP2, s);
  launch_grad_skip(lv2, GA5, GP2, E2, GE2, s);
  launch_conv3x3(2, lv2, GE2, w.wf[2], nullptr, nullptr, GP1, s);
  launch_grad_skip(lv1, GA6, GP1, E1, GE1, s);
  LAUNCH_CHECK("network backward data");
  // weight gradients: each layer's pre-activation gradient against its input
  const struct {
    int l;
    Img m;
    const __bf16 *G, *X;
  } layers[NET_CONVS] = {{6, lv1, GU1, A6}, {5, lv2, GU2, A5}, {4, lv3, GB3, E3},
                         {3, lv3, GE3, P2}, {2, lv2, GE2, P1}, {1, lv1, GE1, XSin}};
  for (const auto& L : layers) {
    const int nb = launch_conv_wgrad(L.m, L.G, L.X, part, s);
    launch_reduce_conv_grad(nb, part, grad + np_w(L.l), grad + np_b(L.l), s);
  }
  const int nb = launch_head_wgrad(lv1, G4, U1, part, s);
  launch_reduce_head_grad(nb, part, grad, s);
  LAUNCH_CHECK("network weight gradients");
  return CDIFF_OK;
}

int cdiff_net::adam(float lr, hipStream_t s) {
  ++adam_step;
  launch_adam(params, grad, m1, m2, NP_TOTAL, lr, 0.9f, 0.999f, 1e-8f, adam_step, s);
  launch_pack_weights(params, w, s);
  LAUNCH_CHECK("adam");
  return CDIFF_OK;
}

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

int cdiff_net_num_params(void) { return NP_TOTAL; }

int cdiff_net_create(uint64_t seed, cdiff_net** out) {
  if (!out) return CDIFF_BAD_ARG;
  cdiff_net* n = new cdiff_net();
  int rc = n->init(seed);
  *out = n;
  return rc;
}

int cdiff_net_destroy(cdiff_net* n) {
  if (!n) return CDIFF_OK;
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  for (void* p : n->ws) (void)hipFree(p);
  for (void* p : n->owned) (void)hipFree(p);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
  return CDIFF_OK;
}

int cdiff_net_set_params(cdiff_net* n, const float* p, int count) {
  if (!n || !p || count != NP_TOTAL) return CDIFF_BAD_ARG;
  hipError_t e = hipMemcpyAsync(n->params, p, sizeof(float) * NP_TOTAL, hipMemcpyHostToDevice, n->stream);
  if (e != hipSuccess) return n->hip_fail(e, "hipMemcpyAsync(params)");
  launch_pack_weights(n->params, n->w, n->stream);
  e = hipStreamSynchronize(n->stream);
  if (e != hipSuccess) return n->hip_fail(e, "set_params");
  n->adam_step = 0;
  (void)hipMemsetAsync(n->m1, 0, sizeof(float) * NP_TOTAL, n->stream);
  (void)hipMemsetAsync(n->m2, 0, sizeof(float) * NP_TOTAL, n->stream);
  e = hipStreamSynchronize(n->stream);
  return e == hipSuccess ? CDIFF_OK : n->hip_fail(e, "set_params");
}

int cdiff_net_get_params(cdiff_net* n, float* p, int count) {
  if (!n || !p || count != NP_TOTAL) return CDIFF_BAD_ARG;
  hipError_t e = hipMemcpyAsync(p, n->params, sizeof(float) * NP_TOTAL, hipMemcpyDeviceToHost, n->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
  return e == hipSuccess ? CDIFF_OK : n->hip_fail(e, "get_params");
}

int cdiff_net_dataset(cdiff_net* n, int height, int width, int capacity) {
  if (!n || height < 1 || width < 1 || capacity < 1) return CDIFF_BAD_ARG;
  if (n->dfeat) return n->fail(CDIFF_BAD_ARG, "the training store is already allocated");
  n->dimg = Img{height, width};
  n->capacity = capacity;
  const size_t fbytes = sizeof(__bf16) * (size_t)n->feat_stride() * capacity;
  const size_t tbytes = sizeof(float) * (size_t)height * width * capacity;
  for (auto pr : {std::make_pair((void**)&n->dfeat, fbytes), std::make_pair((void**)&n->dtgt, tbytes),
                  std::make_pair((void**)&n->dts, 2 * sizeof(double) * capacity),
                  std::make_pair((void**)&n->dloss, sizeof(double) * capacity)}) {
    hipError_t e = hipMalloc(pr.first, pr.second);
    if (e != hipSuccess) return n->hip_fail(e, "hipMalloc(training store)");
    n->owned.push_back(*pr.first);
    e = hipMemsetAsync(*pr.first, 0, pr.second, n->stream);
    if (e != hipSuccess) return n->hip_fail(e, "hipMemsetAsync(training store)");
  }
  hipError_t e = hipStreamSynchronize(n->stream);
  return e == hipSuccess ? CDIFF_OK : n->hip_fail(e, "dataset");
}

int cdiff_net_capture(cdiff_net* n, cdiff_solver* h, int slot, int stage, double dt) {
  if (!n || !h || slot < 0 || slot >= n->capacity || (stage != 0 && stage != 1)) return CDIFF_BAD_ARG;
  if (h->ndim != 2 || h->g.hi[1] - h->g.lo[1] + 1 != n->dimg.H || h->g.hi[0] - h->g.lo[0] + 1 != n->dimg.W)
    return n->fail(CDIFF_BAD_ARG, "the solver's domain does not match the training store");
  int rc;
  if (stage == 0)
    rc = h->learn_features(dt, n->dfeat + (int64_t)slot * n->feat_stride(), n->dts + 2 * slot);
  else
    rc = h->learn_target(n->dts + 2 * slot, n->dtgt + (int64_t)slot * n->dimg.H * n->dimg.W);
  if (rc) return n->fail(rc, h->err);
  hipError_t e = hipStreamSynchronize(h->stream);
  return e == hipSuccess ? CDIFF_OK : n->hip_fail(e, "capture");
}

int cdiff_net_train(cdiff_net* n, int nsamples, int epochs, float lr, uint64_t seed, double* loss) {
  if (!n || nsamples < 1 || nsamples > n->capacity || epochs < 1 || !(lr > 0)) return CDIFF_BAD_ARG;
  int rc = n->workspace(n->dimg);
  if (rc) return rc;
  std::vector<int> order(nsamples);
  std::iota(order.begin(), order.end(), 0);
  std::mt19937_64 rng(seed);
  std::vector<double> l(nsamples);
  for (int ep = 0; ep < epochs; ++ep) {
    std::shuffle(order.begin(), order.end(), rng);
    for (int k = 0; k < nsamples; ++k) {
      const int s = order[k];
      if ((rc = n->backward(n->dfeat + (int64_t)s * n->feat_stride(), n->dtgt + (int64_t)s * n->dimg.H * n->dimg.W,
                            n->dloss + s, n->stream)))
        return rc;
      if ((rc = n->adam(lr, n->stream))) return rc;
    }
  }
  hipError_t e = hipMemcpyAsync(l.data(), n->dloss, sizeof(double) * nsamples, hipMemcpyDeviceToHost, n->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
  if (e != hipSuccess) return n->hip_fail(e, "train");
  if (loss) {
    double s = 0.0;
    for (double v : l) s += v;
    *loss = s / nsamples;
  }
  return CDIFF_OK;
}

int cdiff_net_eval(cdiff_net* n, int slot, float* delta, double* loss) {
  if (!n || slot < 0 || slot >= n->capacity) return CDIFF_BAD_ARG;
  int rc = n->workspace(n->dimg);
  if (rc) return rc;
  const int64_t ni = (int64_t)n->dimg.H * n->dimg.W;
  // the backward pass also gives the loss; its gradient is not applied
  if ((rc = n->backward(n->dfeat + (int64_t)slot * n->feat_stride(), n->dtgt + (int64_t)slot * ni, n->dloss + slot,
                        n->stream)))
    return rc;
  double l = 0.0;
  hipError_t e = hipMemcpyAsync(&l, n->dloss + slot, sizeof(double), hipMemcpyDeviceToHost, n->stream);
  if (e == hipSuccess && delta)
    e = hipMemcpyAsync(delta, n->delta, sizeof(float) * ni, hipMemcpyDeviceToHost, n->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
  if (e != hipSuccess) return n->hip_fail(e, "eval");
  if (loss) *loss = l;
  return CDIFF_OK;
}

int cdiff_net_loss_grad(cdiff_net* n, int height, int width, const uint16_t* features, const float* target,
                        float* grad, float* delta, double* loss) {
  if (!n || !features || !target || height < 1 || width < 1) return CDIFF_BAD_ARG;
  const Img m{height, width};
  int rc = n->workspace(m);
  if (rc) return rc;
  const int64_t ni = (int64_t)height * width;
  hipError_t e = hipMemcpyAsync(n->X0, features, sizeof(uint16_t) * img_pixels(m) * NET_C0, hipMemcpyHostToDevice,
                                n->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(n->tgt, target, sizeof(float) * ni, hipMemcpyHostToDevice, n->stream);
  if (e != hipSuccess) return n->hip_fail(e, "upload");
  launch_s2d(m, n->X0, n->XS, n->stream);
  double* lsum = nullptr;
  e = hipMalloc((void**)&lsum, sizeof(double));
  if (e != hipSuccess) return n->hip_fail(e, "hipMalloc");
  rc = n->backward(n->XS, n->tgt, lsum, n->stream);
  double l = 0.0;
  if (!rc) {
    e = hipMemcpyAsync(&l, lsum, sizeof(double), hipMemcpyDeviceToHost, n->stream);
    if (e == hipSuccess && grad) e = hipMemcpyAsync(grad, n->grad, sizeof(float) * NP_TOTAL, hipMemcpyDeviceToHost, n->stream);
    if (e == hipSuccess && delta) e = hipMemcpyAsync(delta, n->delta, sizeof(float) * ni, hipMemcpyDeviceToHost, n->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
    if (e != hipSuccess) rc = n->hip_fail(e, "download");
  }
  (void)hipFree(lsum);
  if (!rc && loss) *loss = l;
  return rc;
}

int cdiff_net_predict(cdiff_net* n, int height, int width, const uint16_t* features, float* delta) {
  if (!n || !features || !delta || height < 1 || width < 1) return CDIFF_BAD_ARG;
  const Img m{height, width};
  int rc = n->workspace(m);
  if (rc) return rc;
  const int64_t ni = (int64_t)height * width;
  hipError_t e = hipMemcpyAsync(n->X0, features, sizeof(uint16_t) * img_pixels(m) * NET_C0, hipMemcpyHostToDevice,
                                n->stream);
  if (e != hipSuccess) return n->hip_fail(e, "upload");
  launch_s2d(m, n->X0, n->XS, n->stream);
  if ((rc = n->forward(n->XS, n->stream))) return rc;
  e = hipMemcpyAsync(delta, n->delta, sizeof(float) * ni, hipMemcpyDeviceToHost, n->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
  if (e != hipSuccess) return n->hip_fail(e, "download");
  return CDIFF_OK;
}

const char* cdiff_net_last_error(const cdiff_net* n) { return n ? n->err.c_str() : "null network"; }

int cdiff_net_set_loss(cdiff_net* n, int kind) {
  if (!n || (kind != NET_LOSS_MSE && kind != NET_LOSS_RESIDUAL)) return CDIFF_BAD_ARG;
  n->loss_kind = kind;
  return CDIFF_OK;
}

int cdiff_net_set_gate(cdiff_net* n, float gate) {
  if (!n || !(gate >= 0.f)) return CDIFF_BAD_ARG;
  n->gate = gate;
  return CDIFF_OK;
}

int cdiff_set_warm_start(cdiff_solver* h, cdiff_net* n) {
  if (!h) return CDIFF_BAD_ARG;
  h->net = n;
  h->warm_rejects = 0;
  h->warm_skip = 0;
  h->warm_decades = 1.0;
  return CDIFF_OK;
}

}  // extern "C"
