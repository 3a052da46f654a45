// net.h -- the learned warm start's network object (struct cdiff_net).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/cdiff.h"
#include "learn.h"

struct cdiff_net {
  hipStream_t stream = nullptr;  // training stream (inference runs on the solver's)
  float* params = nullptr;       // FP32 master weights (cdiff::NP_TOTAL)
  float* grad = nullptr;
  float* m1 = nullptr;
  float* m2 = nullptr;
  int adam_step = 0;
  float gate = 1e-5f;  // activity threshold of the prediction (relative explicit increment)
  cdiff::NetWeights w{};

  // workspace for one image size
  cdiff::Img wimg{0, 0};
  __bf16 *X0 = nullptr, *X1 = nullptr, *X2 = nullptr, *X3 = nullptr;
  __bf16 *G1 = nullptr, *G2 = nullptr, *G3 = nullptr;
  float *delta = nullptr, *G4 = nullptr, *tgt = nullptr;
  double* loss_part = nullptr;
  float* part = nullptr;
  std::vector<void*> ws;

  // training store
  cdiff::Img dimg{0, 0};
  int capacity = 0;
  __bf16* dfeat = nullptr;
  float* dtgt = nullptr;
  double* dts = nullptr;
  double* dloss = nullptr;  // per-sample loss of the current epoch
  std::vector<void*> owned;
  std::string err;

  int fail(int code, const std::string& msg);
  int hip_fail(hipError_t e, const char* what);
  int init(uint64_t seed);
  int workspace(cdiff::Img m);
  int64_t feat_stride() const { return cdiff::img_pixels(dimg) * cdiff::NET_C0; }
  // delta = net(X0) on `s` (workspace sized for m)
  int forward(const __bf16* X0in, hipStream_t s);
  // loss, gradient into `grad`, per-sample loss into *loss_dev (device)
  int backward(const __bf16* X0in, const float* target, double* loss_dev, hipStream_t s);
  int adam(float lr, hipStream_t s);
};
