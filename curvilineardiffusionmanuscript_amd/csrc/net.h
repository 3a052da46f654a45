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
  int loss_kind = cdiff::NET_LOSS_MSE;  // training loss (learn.h, launch_loss)
  cdiff::NetWeights w{};

  // workspace for one image size (level 0: the domain cells; levels 1-3)
  cdiff::Img wimg{0, 0};
  __bf16* X0 = nullptr;  // level 0, 8 channels: host features before their block layout
  __bf16* XS = nullptr;  // level 1: the network's input
  __bf16 *E1 = nullptr, *A6 = nullptr, *U1 = nullptr;                 // level 1
  __bf16 *P1 = nullptr, *E2 = nullptr, *A5 = nullptr, *U2 = nullptr;  // level 2
  __bf16 *P2 = nullptr, *E3 = nullptr, *B3 = nullptr;                 // level 3
  __bf16 *GU1 = nullptr, *GA6 = nullptr, *GE1 = nullptr;              // gradients, level 1
  __bf16 *GU2 = nullptr, *GA5 = nullptr, *GE2 = nullptr, *GP1 = nullptr;  // level 2
  __bf16 *GB3 = nullptr, *GE3 = nullptr, *GP2 = nullptr;              // level 3
  float *delta = nullptr, *G4 = nullptr, *tgt = nullptr, *res = nullptr;
  double* loss_part = nullptr;
  double* loss_sums = nullptr;
  float* part = nullptr;
  std::vector<void*> ws;

  // training store (features in the level-1 block layout)
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
  int64_t feat_stride() const { return cdiff::img_pixels(cdiff::img_half(dimg)) * cdiff::NET_C; }
  // delta = net(XS) on `s` (workspace sized for the level-0 image)
  int forward(const __bf16* XSin, hipStream_t s);
  // loss, gradient into `grad`, per-sample loss into *loss_dev (device)
  int backward(const __bf16* XSin, const float* target, double* loss_dev, hipStream_t s);
  int adam(float lr, hipStream_t s);
};
