// learn.h -- the learned warm start: a 4-layer 3x3 convolutional network in
// BF16 with FP32 accumulation on MFMA, its training kernels, and the
// feature/target/prediction kernels that connect it to the solver.
//
// Network (per pixel = per domain cell of a 2D mesh):
//   X0 (8 features) -conv3x3-> X1 (32, ReLU) -conv3x3-> X2 (32, ReLU)
//      -conv3x3-> X3 (32, ReLU) -conv3x3-> delta (1, linear)
// delta predicts (T^{n+1} - T^n) / E, E = dt max |r(T^n)| (the largest
// explicit-Euler increment).
//
// Images are BF16, channels-last, with a one-pixel zero border: pixel (y, x)
// of an H x W image (y in [-1, H], x in [-1, W]) is at
// ((y + 1) (W + 2) + x + 1) * C.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cdiff_dev.h"

namespace cdiff {

constexpr int NET_C0 = 8;   // input features (6 used, 2 zero)
constexpr int NET_C = 32;   // hidden channels
// parameter vector (FP32 master copy), in this order
constexpr int NP_W1 = 0, NP_B1 = NP_W1 + NET_C * 9 * NET_C0;  // W1[co][tap][ci]
constexpr int NP_W2 = NP_B1 + NET_C, NP_B2 = NP_W2 + NET_C * 9 * NET_C;
constexpr int NP_W3 = NP_B2 + NET_C, NP_B3 = NP_W3 + NET_C * 9 * NET_C;
constexpr int NP_W4 = NP_B3 + NET_C, NP_B4 = NP_W4 + 9 * NET_C;  // W4[tap][ci]
constexpr int NP_TOTAL = NP_B4 + 1;

struct Img {
  int H, W;
};

__host__ __device__ inline int64_t img_pix(const Img& m, int y, int x) {
  return (int64_t)(y + 1) * (m.W + 2) + (x + 1);
}
inline int64_t img_pixels(const Img& m) { return (int64_t)(m.H + 2) * (m.W + 2); }

// BF16 weight copies used by the kernels
struct NetWeights {
  __bf16* w1p;  // [32][3*32]   W1 packed as 3 K-blocks of 4 taps x 8 channels (tap 9..11 zero)
  __bf16* w2p;  // [32][9*32]   W2 (= master layout)
  __bf16* w3p;
  __bf16* w2f;  // [32][9*32]   flipped/transposed for the backward data pass: [ci][8-tap][co]
  __bf16* w3f;
  __bf16* w4;   // [9][32]
  const float* params;  // FP32 master (biases are read from here)
};

// conv3x3 with 32 output channels on MFMA.  mode 0: Y = relu(conv + bias);
// mode 1: Y = conv * (mask > 0) (backward data pass with the ReLU derivative).
void launch_conv3x3(int cin, int mode, Img m, const __bf16* X, const __bf16* Wp, const float* bias,
                    const __bf16* mask, __bf16* Y, hipStream_t s);
// head: delta = conv3x3(X3, W4) + b4 (FP32 out, [H][W]) on active pixels
// (|feature 5| > gate in the 3x3 neighbourhood of X0), 0 elsewhere.  With
// target: also g = 2 (delta - target) / (H W) on active pixels into G4 (FP32,
// bordered image) and per-block partial sums of (delta - target)^2.
// inference head on MFMA (no loss); matches launch_head_fwd up to FP32 summation order
void launch_head_infer(Img m, const __bf16* X3, const __bf16* w4, const float* params, const __bf16* X0, float gate,
                       float* delta, hipStream_t s);
void launch_head_fwd(Img m, const __bf16* X3, const __bf16* w4, const float* params, const __bf16* X0, float gate,
                     float* delta, const float* target, float* G4, double* loss_part, int* nparts, hipStream_t s);
// G3pre = (sum_tap G4(p - off) W4[tap]) * (X3 > 0)
void launch_head_bwd(Img m, const float* G4, const __bf16* w4, const __bf16* X3, __bf16* G3, hipStream_t s);
// weight (and bias) gradient partials of a conv layer: one row of 32 x npad
// per block; returns the number of blocks
int launch_conv_wgrad(int cin, Img m, const __bf16* G, const __bf16* X, float* part, hipStream_t s);
int conv_wgrad_npad(int cin);
// head weight gradient partials: 289 per block
int launch_head_wgrad(Img m, const float* G4, const __bf16* X3, float* part, hipStream_t s);
// sum of partial rows into the gradient vector (fixed order)
void launch_reduce_conv_grad(int cin, int nblocks, const float* part, float* gW, float* gB, hipStream_t s);
void launch_reduce_head_grad(int nblocks, const float* part, float* grad, hipStream_t s);
// Adam on the FP32 master, then the BF16 copies
void launch_adam(float* params, const float* grad, float* m1, float* m2, int n, float lr, float b1, float b2,
                 float eps, int step, hipStream_t s);
void launch_pack_weights(const float* params, NetWeights w, hipStream_t s);
void launch_fill_bf16(__bf16* p, int64_t n, hipStream_t s);
void launch_sum_loss(const double* part, int n, double* out, hipStream_t s);

// Solver side (2D meshes).
struct FeatIn {
  Geo g;
  int harmonic;
  double dt, dt_ratio, j_ref;
  const double* u;    // T^n (full)
  const double* um1;  // T^{n-1} (full)
  const double* a;    // alpha(T^n) (full)
  const double* g11;
  const double* g22;
  const double* jac;  // cell Jacobian
  const float* r32;   // FP64 residual of T^n, stored FP32
  const double* scales;  // device: [0] max |T^n|, [1] dt max |r(T^n)| over the domain
};
void launch_features(const FeatIn& f, __bf16* X0, hipStream_t s);
// target (T^{n+1} - T^n) / scales[1]; prediction T^n + scales[1] delta
void launch_target(Geo g, const double* unew, const double* uold, const double* scales, float* tgt, hipStream_t s);
void launch_apply_prediction(Geo g, const double* u, const float* delta, const double* scales, double* out,
                             hipStream_t s);
// max |u| (FP64) or mul * max |r| (FP32) over the domain into *out (two passes;
// part holds >= 1024 doubles)
void launch_domain_maxabs(Geo g, const double* u, double* part, double* out, hipStream_t s);
void launch_domain_maxabs32(Geo g, const float* r, double mul, double* part, double* out, hipStream_t s);

}  // namespace cdiff
