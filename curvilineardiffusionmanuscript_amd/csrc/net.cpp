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
  HIP_TRY(alloc((void**)&w.w1p, sizeof(__bf16) * NET_C * 96));
  HIP_TRY(alloc((void**)&w.w2p, sizeof(__bf16) * NET_C * 9 * NET_C));
  HIP_TRY(alloc((void**)&w.w3p, sizeof(__bf16) * NET_C * 9 * NET_C));
  HIP_TRY(alloc((void**)&w.w2f, sizeof(__bf16) * NET_C * 9 * NET_C));
  HIP_TRY(alloc((void**)&w.w3f, sizeof(__bf16) * NET_C * 9 * NET_C));
  HIP_TRY(alloc((void**)&w.w4, sizeof(__bf16) * 9 * NET_C));
  w.params = params;
  // He initialisation of the ReLU layers, a small head: the first predictions
  // are near zero (the warm start then falls back to T^n)
  std::vector<float> p(NP_TOTAL, 0.f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.f, 1.f);
  auto fill = [&](int off, int n, float sd) {
    for (int i = 0; i < n; ++i) p[off + i] = sd * nd(rng);
  };
  fill(NP_W1, NET_C * 9 * NET_C0, std::sqrt(2.f / (9 * 6)));
  fill(NP_W2, NET_C * 9 * NET_C, std::sqrt(2.f / (9 * NET_C)));
  fill(NP_W3, NET_C * 9 * NET_C, std::sqrt(2.f / (9 * NET_C)));
  fill(NP_W4, 9 * NET_C, 0.01f * std::sqrt(1.f / (9 * NET_C)));
  HIP_TRY(hipMemcpyAsync(params, p.data(), sizeof(float) * NP_TOTAL, hipMemcpyHostToDevice, stream));
  launch_pack_weights(params, w, stream);
  LAUNCH_CHECK("k_pack_weights");
  HIP_TRY(hipStreamSynchronize(stream));
  return CDIFF_OK;
}

int cdiff_net::workspace(Img m) {
  if (m.H == wimg.H && m.W == wimg.W && X1) return CDIFF_OK;
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
  HIP_TRY(alloc((void**)&X0, sizeof(__bf16) * npx * NET_C0));
  for (__bf16** b : {&X1, &X2, &X3, &G1, &G2, &G3}) HIP_TRY(alloc((void**)b, sizeof(__bf16) * npx * NET_C));
  HIP_TRY(alloc((void**)&delta, sizeof(float) * ni));
  HIP_TRY(alloc((void**)&tgt, sizeof(float) * ni));
  HIP_TRY(alloc((void**)&G4, sizeof(float) * npx));
  HIP_TRY(alloc((void**)&loss_part, sizeof(double) * ((ni + 255) / 256)));
  // weight-gradient partials: the largest of the conv layers and the head
  const int64_t conv_blocks = (int64_t)((m.W + 63) / 64) * ((m.H + 31) / 32);
  const int64_t conv_part = conv_blocks * NET_C * conv_wgrad_npad(NET_C);
  const int64_t head_part = ((ni + 4095) / 4096) * (9 * NET_C + 1);
  HIP_TRY(alloc((void**)&part, sizeof(float) * std::max(conv_part, head_part)));
  HIP_TRY(hipStreamSynchronize(stream));
  wimg = m;
  return CDIFF_OK;
}

int cdiff_net::forward(const __bf16* X0in, hipStream_t s) {
  const Img m = wimg;
  launch_conv3x3(NET_C0, 0, m, X0in, w.w1p, params + NP_B1, nullptr, X1, s);
  launch_conv3x3(NET_C, 0, m, X1, w.w2p, params + NP_B2, nullptr, X2, s);
  launch_conv3x3(NET_C, 0, m, X2, w.w3p, params + NP_B3, nullptr, X3, s);
  launch_head_infer(m, X3, w.w4, params, X0in, gate, delta, s);
  LAUNCH_CHECK("network forward");
  return CDIFF_OK;
}

int cdiff_net::backward(const __bf16* X0in, const float* target, double* loss_dev, hipStream_t s) {
  const Img m = wimg;
  launch_conv3x3(NET_C0, 0, m, X0in, w.w1p, params + NP_B1, nullptr, X1, s);
  launch_conv3x3(NET_C, 0, m, X1, w.w2p, params + NP_B2, nullptr, X2, s);
  launch_conv3x3(NET_C, 0, m, X2, w.w3p, params + NP_B3, nullptr, X3, s);
  int nparts = 0;
  launch_head_fwd(m, X3, w.w4, params, X0in, gate, delta, target, G4, loss_part, &nparts, s);
  launch_sum_loss(loss_part, nparts, loss_dev, s);
  // backward data: head, layer 3, layer 2 (with the ReLU derivatives)
  launch_head_bwd(m, G4, w.w4, X3, G3, s);
  launch_conv3x3(NET_C, 1, m, G3, w.w3f, nullptr, X2, G2, s);
  launch_conv3x3(NET_C, 1, m, G2, w.w2f, nullptr, X1, G1, s);
  LAUNCH_CHECK("network backward data");
  // weight gradients
  int nb = launch_conv_wgrad(NET_C, m, G3, X2, part, s);
  launch_reduce_conv_grad(NET_C, nb, part, grad + NP_W3, grad + NP_B3, s);
  nb = launch_conv_wgrad(NET_C, m, G2, X1, part, s);
  launch_reduce_conv_grad(NET_C, nb, part, grad + NP_W2, grad + NP_B2, s);
  nb = launch_conv_wgrad(NET_C0, m, G1, X0in, part, s);
  launch_reduce_conv_grad(NET_C0, nb, part, grad + NP_W1, grad + NP_B1, s);
  nb = launch_head_wgrad(m, G4, X3, part, s);
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
    const double npx = (double)n->dimg.H * n->dimg.W;
    double s = 0.0;
    for (double v : l) s += v / npx;
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
  if (loss) *loss = l / (double)ni;
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
  double* lsum = nullptr;
  e = hipMalloc((void**)&lsum, sizeof(double));
  if (e != hipSuccess) return n->hip_fail(e, "hipMalloc");
  rc = n->backward(n->X0, n->tgt, lsum, n->stream);
  double l = 0.0;
  if (!rc) {
    e = hipMemcpyAsync(&l, lsum, sizeof(double), hipMemcpyDeviceToHost, n->stream);
    if (e == hipSuccess && grad) e = hipMemcpyAsync(grad, n->grad, sizeof(float) * NP_TOTAL, hipMemcpyDeviceToHost, n->stream);
    if (e == hipSuccess && delta) e = hipMemcpyAsync(delta, n->delta, sizeof(float) * ni, hipMemcpyDeviceToHost, n->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
    if (e != hipSuccess) rc = n->hip_fail(e, "download");
  }
  (void)hipFree(lsum);
  if (!rc && loss) *loss = l / (double)ni;
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
  if ((rc = n->forward(n->X0, n->stream))) return rc;
  e = hipMemcpyAsync(delta, n->delta, sizeof(float) * ni, hipMemcpyDeviceToHost, n->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
  if (e != hipSuccess) return n->hip_fail(e, "download");
  return CDIFF_OK;
}

const char* cdiff_net_last_error(const cdiff_net* n) { return n ? n->err.c_str() : "null network"; }

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
