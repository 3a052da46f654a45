// learn.h -- the learned warm start: a 4-level convolutional encoder-decoder
// in BF16 with FP32 accumulation on MFMA, its training kernels, and the
// feature/target/prediction kernels that connect it to the solver.
//
// Network.  Level 0 is the H x W grid of domain cells; level l = 1..3 has
// H_l = ceil(H_{l-1} / 2) rows (same for columns).  All layers carry 32
// channels:
//   XS = S2D(X0)                  level 1: the 8 features of each 2x2 block
//                                 of cells, channel 8 (2 dy + dx) + f
//   E1 = relu(conv1(XS))          level 1
//   E2 = relu(conv2(pool(E1)))    level 2     pool: mean of a 2x2 block
//   E3 = relu(conv3(pool(E2)))    level 3
//   B3 = relu(conv4(E3))          level 3
//   U2 = relu(conv5(up(B3) + E2)) level 2     up: each pixel to its 2x2 block
//   U1 = relu(conv6(up(U2) + E1)) level 1
//   delta(2y + dy, 2x + dx) = head(U1)(y, x)[2 dy + dx]   (conv 32 -> 4)
// conv1..conv6 are 3x3 convolutions 32 -> 32 with bias; the head a 3x3
// convolution 32 -> 4 with bias.  delta predicts (T^{n+1} - T^n) / E,
// E = dt max |r(T^n)| (the largest explicit-Euler increment), and is 0 where
// the 3x3 neighbourhood of a cell shows no activity (feature 7, the largest
// |feature 5| there, <= gate).
//
// Images are BF16 (or FP32 for the head's gradient), channels-last, with a
// one-pixel zero border: pixel (y, x) of an H x W image (y in [-1, H], x in
// [-1, W]) is at ((y + 1) (W + 2) + x + 1) * C.  Pixels of a block beyond an
// odd image edge are 0.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cdiff_dev.h"

namespace cdiff {

constexpr int NET_C0 = 8;   // features per cell (learn.hip, k_features)
constexpr int NET_C = 32;   // channels of every layer
constexpr int NET_HO = 4;   // head outputs per level-1 pixel (its 2x2 block of cells)
constexpr int NET_CONVS = 6;
constexpr int NP_CONV = NET_C * 9 * NET_C + NET_C;  // W[co][tap][ci], then b[co]
// parameter vector (FP32 master copy): conv1..conv6, then the head
constexpr int np_w(int l) { return (l - 1) * NP_CONV; }  // l = 1..6
constexpr int np_b(int l) { return (l - 1) * NP_CONV + NET_C * 9 * NET_C; }
constexpr int NP_WH = NET_CONVS * NP_CONV;         // head W[s][tap][ci]
constexpr int NP_BH = NP_WH + NET_HO * 9 * NET_C;  // head b[s]
constexpr int NP_TOTAL = NP_BH + NET_HO;

struct Img {
  int H, W;
};

__host__ __device__ inline int64_t img_pix(const Img& m, int y, int x) {
  return (int64_t)(y + 1) * (m.W + 2) + (x + 1);
}
inline int64_t img_pixels(const Img& m) { return (int64_t)(m.H + 2) * (m.W + 2); }
inline Img img_half(const Img& m) { return Img{(m.H + 1) / 2, (m.W + 1) / 2}; }

// BF16 weight copies used by the kernels
struct NetWeights {
  __bf16* wp[NET_CONVS + 1];  // [1..6]: [32][9*32] (= master layout)
  __bf16* wf[NET_CONVS + 1];  // [2..6]: flipped/transposed for the backward data pass: [ci][8-tap][co]
  __bf16* wh;                 // head [4][9][32]
  const float* params;        // FP32 master (biases are read from here)
};

// conv3x3 32 -> 32 on MFMA.  mode 0: Y = relu(conv + bias); mode 1:
// Y = conv * (mask > 0) (backward data with the ReLU derivative of the layer
// below); mode 2: Y = conv (backward data, the derivative applied later).
void launch_conv3x3(int mode, Img m, const __bf16* X, const __bf16* Wp, const float* bias, const __bf16* mask,
                    __bf16* Y, hipStream_t s);
// P (level l + 1) = mean of the 2x2 blocks of X (level l of size m)
void launch_pool(Img m, const __bf16* X, __bf16* P, hipStream_t s);
// A (level l, size m) = up(C) + S, C of level l + 1
void launch_up_add(Img m, const __bf16* C, const __bf16* S, __bf16* A, hipStream_t s);
// GE (size m) = (GA + pool^T(GP)) * (E > 0): the skip's gradient plus the
// pooled branch's, GP of level l + 1 (pool^T spreads a quarter to each pixel)
void launch_grad_skip(Img m, const __bf16* GA, const __bf16* GP, const __bf16* E, __bf16* GE, hipStream_t s);
// GU (level l + 1) = up^T(GA) * (U > 0), GA of level l (size m): the sum over
// each 2x2 block
void launch_grad_up(Img m, const __bf16* GA, const __bf16* U, __bf16* GU, hipStream_t s);
// X0 (level 0, 8 channels) -> XS (level 1, 32 channels), feature 7 formed
// from the 3x3 neighbourhoods of feature 5
void launch_s2d(Img m0, const __bf16* X0, __bf16* XS, hipStream_t s);
// head (inference, MFMA): delta (FP32, [H][W] of level 0 m0) from U1, gated by
// feature 7 of XS
void launch_head_infer(Img m0, const __bf16* U1, const __bf16* wh, const float* params, const __bf16* XS, float gate,
                       float* delta, hipStream_t s);
// head (training, vector ALU): delta
void launch_head_fwd(Img m0, const __bf16* U1, const __bf16* wh, const float* params, const __bf16* XS, float gate,
                     float* delta, hipStream_t s);
// The training loss into *loss and its gradient G4 (FP32, level 1 bordered,
// 4 channels, 0 on inactive cells).  kind 0: the mean squared error of the
// increment, |e|^2 / (H W), e = delta - target; kind 1: |A e|^2 / |A target|^2,
// A the implicit step I - dt L with the face diffusion numbers
// expm1(features 3, 4) -- the squared ratio of the residuals of the predicted
// start and of T^n, to first order.  re: H*W floats; part: 2 loss_parts(m0)
// doubles; sums: 2 doubles.
constexpr int NET_LOSS_MSE = 0, NET_LOSS_RESIDUAL = 1;
int loss_parts(Img m0);
void launch_loss(int kind, Img m0, const __bf16* XS, const float* delta, const float* target, float gate, float* re,
                 double* part, double* sums, double* loss, float* G4, hipStream_t s);
// GU1 = head^T(G4) * (U1 > 0)  (level 1 of m0)
void launch_head_bwd(Img m0, const float* G4, const __bf16* wh, const __bf16* U1, __bf16* GU1, hipStream_t s);
// weight (and bias) gradient partials of a conv layer (image m): one row of
// 32 x conv_wgrad_npad() per block; returns the number of blocks
int launch_conv_wgrad(Img m, const __bf16* G, const __bf16* X, float* part, hipStream_t s);
int conv_wgrad_npad();
int conv_wgrad_blocks(Img m);
// head weight gradient partials (level 1 image m1): 4 x 289 per block
int launch_head_wgrad(Img m1, const float* G4, const __bf16* U1, float* part, hipStream_t s);
int head_wgrad_blocks(Img m1);
// sum of partial rows into the gradient vector (fixed order)
void launch_reduce_conv_grad(int nblocks, const float* part, float* gW, float* gB, hipStream_t s);
void launch_reduce_head_grad(int nblocks, const float* part, float* grad, hipStream_t s);
// Adam on the FP32 master, then the BF16 copies
void launch_adam(float* params, const float* grad, float* m1, float* m2, int n, float lr, float b1, float b2,
                 float eps, int step, hipStream_t s);
void launch_pack_weights(const float* params, NetWeights w, hipStream_t s);
void launch_fill_bf16(__bf16* p, int64_t n, hipStream_t s);

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
// the features of every domain cell, written in the level-1 block layout XS
void launch_features(const FeatIn& f, __bf16* XS, hipStream_t s);
// target (T^{n+1} - T^n) / scales[1]; prediction T^n + scales[1] delta
void launch_target(Geo g, const double* unew, const double* uold, const double* scales, float* tgt, hipStream_t s);
void launch_apply_prediction(Geo g, const double* u, const float* delta, const double* scales, double* out,
                             hipStream_t s);
// max |u| (FP64) or mul * max |r| (FP32) over the domain into *out (two passes;
// part holds >= 1024 doubles)
void launch_domain_maxabs(Geo g, const double* u, double* part, double* out, hipStream_t s);
void launch_domain_maxabs32(Geo g, const float* r, double mul, double* part, double* out, hipStream_t s);

}  // namespace cdiff
