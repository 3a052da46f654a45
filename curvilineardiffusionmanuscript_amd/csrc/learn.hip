// learn.hip -- kernels of the learned warm start (see learn.h).
//
// The seven convolutions run on MFMA (v_mfma_f32_16x16x32_bf16, BF16
// operands, FP32 accumulation) as implicit GEMMs:
//   forward / backward-data   D[co][px] = sum_k W[co][k] * X[px + off(k)][ci(k)]
//                             k = (tap, ci): one 32-wide K-block per tap (a
//                             lane's 8 K-values are 8 channels of one pixel:
//                             one 16-byte load); 16 pixels per N-tile
//   weight gradient           D[co][(tap, ci)] = sum_px G[px][co] X[px + off][ci]
//                             K = pixels, staged in LDS per 4 x 64 tile; the
//                             bias gradient is one more column with X = 1
// MFMA operand layout (16x16x32): lane l holds A[l%16][8(l/16) + j],
// B[8(l/16) + j][l%16] (j = 0..7) and receives D[4(l/16) + r][l%16].
// The 4-output head (1152 MACs per level-1 pixel) runs on MFMA for inference
// (one M-tile, rows 0-3 real) and on the vector ALUs in training, where it
// also forms the loss gradient.  Pooling, up-sampling, the skip additions and
// their gradients are elementwise passes over 8-channel (16-byte) groups.
#include <cmath>

#include "learn.h"

namespace cdiff {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 load8(const __bf16* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}

__device__ __forceinline__ void store8(__bf16* p, bf16x8 v) {
  *reinterpret_cast<uint4*>(p) = __builtin_bit_cast(uint4, v);
}

__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u)); }

__device__ __forceinline__ float bf(const __bf16 v) { return (float)v; }

constexpr int CONV_RB = 16;     // output rows per forward block
constexpr int CG = NET_C / 8;   // 8-channel groups per pixel

// 3x3 convolution, 32 -> 32 channels, as an implicit GEMM on MFMA.  A block
// of 4 waves owns a 64-pixel-wide strip of CONV_RB rows; wave w computes the
// 16-pixel N-tile w of every row for both 16-channel M-tiles, with the
// weights (A operands) in registers for the whole strip.  The input rows pass
// through a 3-row LDS ring (66 pixels x 32 channels): each input row is read
// from memory once per strip, the next row's load is in flight while the
// current row is computed, and the nine taps read their B operands (8
// channels of one pixel per lane) from LDS.  The MFMA chain per output tile
// is the same as a direct evaluation (taps in order), so results do not
// depend on the staging.
//   MODE 0: y = relu(acc + bias)   MODE 1: y = acc where mask > 0, else 0
//   MODE 2: y = acc
template <int MODE>
__global__ __launch_bounds__(256) void k_conv3x3(Img m, const __bf16* __restrict__ X, const __bf16* __restrict__ Wp,
                                                 const float* __restrict__ bias, const __bf16* __restrict__ mask,
                                                 __bf16* __restrict__ Y) {
  constexpr int KB = 9;                  // K-blocks of 32: one per tap
  constexpr int RX = 66;                 // ring row: pixels x0-1 .. x0+64
  constexpr int NCH = RX * CG;           // 16-byte chunks per ring row
  constexpr int NLD = (NCH + 255) / 256;
  __shared__ bf16x8 sRow[3][NCH];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int x0 = blockIdx.x * 64, y0 = blockIdx.y * CONV_RB;
  const int y1 = min(y0 + CONV_RB, m.H);  // rows y0 .. y1-1
  bf16x8 a[2][KB];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) a[mt][kb] = load8(Wp + (16 * mt + lr) * (KB * 32) + kb * 32 + 8 * lg);

  // input row yi (-1 .. H, the zero border included) into registers
  bf16x8 pre[NLD];
  auto fetch = [&](int yi) {
#pragma unroll
    for (int q = 0; q < NLD; ++q) {
      const int t = tid + q * 256;
      pre[q] = zero8();
      if (t < NCH) {
        const int px = t / CG, c = t % CG;
        const int xi = x0 - 1 + px;
        if (yi <= m.H && xi <= m.W) pre[q] = load8(X + img_pix(m, yi, xi) * NET_C + 8 * c);
      }
    }
  };
  auto store = [&](int yi) {
    bf16x8* R = sRow[(yi + 3) % 3];
#pragma unroll
    for (int q = 0; q < NLD; ++q) {
      const int t = tid + q * 256;
      if (t < NCH) R[t] = pre[q];
    }
  };
  fetch(y0 - 1);
  store(y0 - 1);
  fetch(y0);
  store(y0);
  fetch(y0 + 1);
  const int xo = 16 * w + lr;  // output pixel of this lane within the strip
  for (int y = y0; y < y1; ++y) {
    store(y + 1);
    __syncthreads();  // rows y-1, y, y+1 in LDS
    if (y + 1 < y1) fetch(y + 2);
    f32x4 acc[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      const int dy = kb / 3 - 1, dx = kb % 3 - 1;
      const bf16x8 bv = sRow[(y + dy + 3) % 3][(xo + 1 + dx) * CG + lg];
      acc[0] = mfma16(a[0][kb], bv, acc[0]);
      acc[1] = mfma16(a[1][kb], bv, acc[1]);
    }
    const int x = x0 + xo;
    if (x < m.W) {
      const int64_t p = img_pix(m, y, x);
#pragma unroll
      for (int mt = 0; mt < 2; ++mt) {
        const int co = 16 * mt + 4 * lg;
        bf16x4 o;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          float v = acc[mt][r];
          if (MODE == 0)
            v = fmaxf(v + bias[co + r], 0.f);
          else if (MODE == 1)
            v = bf(mask[p * NET_C + co + r]) > 0.f ? v : 0.f;
          o[r] = (__bf16)v;
        }
        *reinterpret_cast<bf16x4*>(Y + p * NET_C + co) = o;
      }
    }
    __syncthreads();  // row y-1's slot is reloaded next
  }
}

// ---- pooling, up-sampling, skips and their gradients (one thread per
//      8-channel group of a pixel)

// P[y][x] = ((X[2y][2x] + X[2y][2x+1]) + (X[2y+1][2x] + X[2y+1][2x+1])) / 4,
// pixels beyond an odd edge read the zero border
__global__ void k_pool(Img m, Img mc, const __bf16* __restrict__ X, __bf16* __restrict__ P) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)mc.H * mc.W * CG) return;
  const int c = (int)(e % CG);
  const int64_t px = e / CG;
  const int y = (int)(px / mc.W), x = (int)(px % mc.W);
  const bf16x8 a = load8(X + img_pix(m, 2 * y, 2 * x) * NET_C + 8 * c);
  const bf16x8 b = load8(X + img_pix(m, 2 * y, 2 * x + 1) * NET_C + 8 * c);
  const bf16x8 d = load8(X + img_pix(m, 2 * y + 1, 2 * x) * NET_C + 8 * c);
  const bf16x8 f = load8(X + img_pix(m, 2 * y + 1, 2 * x + 1) * NET_C + 8 * c);
  bf16x8 o;
#pragma unroll
  for (int j = 0; j < 8; ++j) o[j] = (__bf16)(((bf(a[j]) + bf(b[j])) + (bf(d[j]) + bf(f[j]))) * 0.25f);
  store8(P + img_pix(mc, y, x) * NET_C + 8 * c, o);
}

// A[y][x] = C[y/2][x/2] + S[y][x]
__global__ void k_up_add(Img m, Img mc, const __bf16* __restrict__ C, const __bf16* __restrict__ S,
                         __bf16* __restrict__ A) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)m.H * m.W * CG) return;
  const int c = (int)(e % CG);
  const int64_t px = e / CG;
  const int y = (int)(px / m.W), x = (int)(px % m.W);
  const int64_t p = img_pix(m, y, x) * NET_C + 8 * c;
  const bf16x8 u = load8(C + img_pix(mc, y >> 1, x >> 1) * NET_C + 8 * c);
  const bf16x8 s = load8(S + p);
  bf16x8 o;
#pragma unroll
  for (int j = 0; j < 8; ++j) o[j] = (__bf16)(bf(u[j]) + bf(s[j]));
  store8(A + p, o);
}

// GE[y][x] = (GA[y][x] + GP[y/2][x/2] / 4) * (E[y][x] > 0)
__global__ void k_grad_skip(Img m, Img mc, const __bf16* __restrict__ GA, const __bf16* __restrict__ GP,
                            const __bf16* __restrict__ E, __bf16* __restrict__ GE) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)m.H * m.W * CG) return;
  const int c = (int)(e % CG);
  const int64_t px = e / CG;
  const int y = (int)(px / m.W), x = (int)(px % m.W);
  const int64_t p = img_pix(m, y, x) * NET_C + 8 * c;
  const bf16x8 ga = load8(GA + p), em = load8(E + p);
  const bf16x8 gp = load8(GP + img_pix(mc, y >> 1, x >> 1) * NET_C + 8 * c);
  bf16x8 o;
#pragma unroll
  for (int j = 0; j < 8; ++j) o[j] = (__bf16)(bf(em[j]) > 0.f ? bf(ga[j]) + 0.25f * bf(gp[j]) : 0.f);
  store8(GE + p, o);
}

// GU[y][x] = ((GA[2y][2x] + GA[2y][2x+1]) + (GA[2y+1][2x] + GA[2y+1][2x+1])) * (U[y][x] > 0)
__global__ void k_grad_up(Img m, Img mc, const __bf16* __restrict__ GA, const __bf16* __restrict__ U,
                          __bf16* __restrict__ GU) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)mc.H * mc.W * CG) return;
  const int c = (int)(e % CG);
  const int64_t px = e / CG;
  const int y = (int)(px / mc.W), x = (int)(px % mc.W);
  const bf16x8 a = load8(GA + img_pix(m, 2 * y, 2 * x) * NET_C + 8 * c);
  const bf16x8 b = load8(GA + img_pix(m, 2 * y, 2 * x + 1) * NET_C + 8 * c);
  const bf16x8 d = load8(GA + img_pix(m, 2 * y + 1, 2 * x) * NET_C + 8 * c);
  const bf16x8 f = load8(GA + img_pix(m, 2 * y + 1, 2 * x + 1) * NET_C + 8 * c);
  const int64_t q = img_pix(mc, y, x) * NET_C + 8 * c;
  const bf16x8 um = load8(U + q);
  bf16x8 o;
#pragma unroll
  for (int j = 0; j < 8; ++j) o[j] = (__bf16)(bf(um[j]) > 0.f ? (bf(a[j]) + bf(b[j])) + (bf(d[j]) + bf(f[j])) : 0.f);
  store8(GU + q, o);
}

// level-0 features (8 channels) into the level-1 block layout; feature 7 is
// the largest |feature 5| over the cell's 3x3 neighbourhood (its activity)
__global__ void k_s2d(Img m0, Img m1, const __bf16* __restrict__ X0, __bf16* __restrict__ XS) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)m0.H * m0.W) return;
  const int y = (int)(e / m0.W), x = (int)(e % m0.W);
  const int s = 2 * (y & 1) + (x & 1);
  bf16x8 v = load8(X0 + img_pix(m0, y, x) * NET_C0);
  float act = 0.f;
  for (int tap = 0; tap < 9; ++tap)
    act = fmaxf(act, fabsf(bf(X0[img_pix(m0, y + tap / 3 - 1, x + tap % 3 - 1) * NET_C0 + 5])));
  v[7] = (__bf16)act;
  store8(XS + img_pix(m1, y >> 1, x >> 1) * NET_C + 8 * s, v);
}

// ---- head

// A cell is active when the explicit-Euler increment (feature 5, scaled by
// its domain maximum) exceeds `gate` in its 3x3 neighbourhood: feature 7 of
// the cell, read from the level-1 block layout; elsewhere the prediction is 0
// (no change), so the network adds no background error.
__device__ __forceinline__ bool active(const Img& m1, const __bf16* __restrict__ XS, int y, int x, float gate) {
  return bf(XS[img_pix(m1, y >> 1, x >> 1) * NET_C + 8 * (2 * (y & 1) + (x & 1)) + 7]) > gate;
}

// head forward for training (vector ALU): one thread per level-1 pixel, its
// four outputs for the cells of its 2x2 block
__global__ __launch_bounds__(256) void k_head_fwd(Img m0, Img m1, const __bf16* __restrict__ U1,
                                                  const __bf16* __restrict__ wh, const float* __restrict__ params,
                                                  const __bf16* __restrict__ XS, float gate,
                                                  float* __restrict__ delta) {
  __shared__ float sw[NET_HO * 9 * NET_C];
  for (int t = threadIdx.x; t < NET_HO * 9 * NET_C; t += 256) sw[t] = bf(wh[t]);
  __syncthreads();
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)m1.H * m1.W) return;
  const int y = (int)(idx / m1.W), x = (int)(idx % m1.W);
  float acc[NET_HO];
#pragma unroll
  for (int s = 0; s < NET_HO; ++s) acc[s] = params[NP_BH + s];
  for (int tap = 0; tap < 9; ++tap) {
    const __bf16* src = U1 + img_pix(m1, y + tap / 3 - 1, x + tap % 3 - 1) * NET_C;
#pragma unroll
    for (int q = 0; q < CG; ++q) {
      const bf16x8 v = load8(src + 8 * q);
#pragma unroll
      for (int j = 0; j < 8; ++j)
#pragma unroll
        for (int s = 0; s < NET_HO; ++s) acc[s] += bf(v[j]) * sw[(s * 9 + tap) * NET_C + 8 * q + j];
    }
  }
#pragma unroll
  for (int s = 0; s < NET_HO; ++s) {
    const int fy = 2 * y + (s >> 1), fx = 2 * x + (s & 1);
    if (fy < m0.H && fx < m0.W) delta[(int64_t)fy * m0.W + fx] = active(m1, XS, fy, fx, gate) ? acc[s] : 0.f;
  }
}

// ---- the training loss: |A e|^2 / |A target|^2, e = delta - target, with
//      (A v)(c) = v(c) + sum_f D_f (v(c) - v(nbr_f)) over the faces of c that
//      have a neighbour in the image, D_f = expm1(feature 3) on the east
//      face, expm1(feature 4) on the north face (the face diffusion numbers)

__device__ __forceinline__ float feat(const Img& m1, const __bf16* __restrict__ XS, int y, int x, int f) {
  return bf(XS[img_pix(m1, y >> 1, x >> 1) * NET_C + 8 * (2 * (y & 1) + (x & 1)) + f]);
}

// (A v)(y, x) for v given cell-wise by `val`
template <class V>
__device__ __forceinline__ float apply_A(const Img& m0, const Img& m1, const __bf16* __restrict__ XS, int y, int x,
                                         V val) {
  const float vc = val(y, x);
  float r = vc;
  if (x + 1 < m0.W) r += expm1f(feat(m1, XS, y, x, 3)) * (vc - val(y, x + 1));
  if (x > 0) r += expm1f(feat(m1, XS, y, x - 1, 3)) * (vc - val(y, x - 1));
  if (y + 1 < m0.H) r += expm1f(feat(m1, XS, y, x, 4)) * (vc - val(y + 1, x));
  if (y > 0) r += expm1f(feat(m1, XS, y - 1, x, 4)) * (vc - val(y - 1, x));
  return r;
}

// re = A e per cell, and per-block partial sums of (A e)^2 and (A target)^2;
// with RESIDUAL false (the mean squared error of the increment) A is the
// identity and (A target)^2 counts 1 per cell
template <bool RESIDUAL>
__global__ __launch_bounds__(256) void k_res_fwd(Img m0, Img m1, const __bf16* __restrict__ XS,
                                                 const float* __restrict__ delta, const float* __restrict__ target,
                                                 float* __restrict__ re, double* __restrict__ part) {
  __shared__ double sred[2][256 / WAVE];
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double se = 0.0, st = 0.0;
  if (idx < (int64_t)m0.H * m0.W) {
    const int y = (int)(idx / m0.W), x = (int)(idx % m0.W);
    float ae, at;
    if (RESIDUAL) {
      ae = apply_A(m0, m1, XS, y, x, [&](int yy, int xx) {
        const int64_t q = (int64_t)yy * m0.W + xx;
        return delta[q] - target[q];
      });
      at = apply_A(m0, m1, XS, y, x, [&](int yy, int xx) { return target[(int64_t)yy * m0.W + xx]; });
    } else {
      ae = delta[idx] - target[idx];
      at = 1.f;
    }
    re[idx] = ae;
    se = (double)ae * ae;
    st = (double)at * at;
  }
  se = wave_sum(se);
  st = wave_sum(st);
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    sred[0][threadIdx.x / WAVE] = se;
    sred[1][threadIdx.x / WAVE] = st;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int i = 0; i < 256 / WAVE; ++i) {
      a += sred[0][i];
      b += sred[1][i];
    }
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

// sums[0] = sum (A e)^2, sums[1] = sum (A target)^2 (fixed order); the loss
__global__ void k_res_sums(const double* __restrict__ part, int n, double* __restrict__ sums,
                           double* __restrict__ loss) {
  __shared__ double s[2][256];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  s[0][threadIdx.x] = a;
  s[1][threadIdx.x] = b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s[0][threadIdx.x] += s[0][threadIdx.x + o];
      s[1][threadIdx.x] += s[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    sums[0] = s[0][0];
    sums[1] = s[1][0];
    *loss = s[1][0] > 0.0 ? s[0][0] / s[1][0] : 0.0;
  }
}

// G4 (level-1 bordered, 4 channels) = 2 (A re) / sums[1] on active cells, 0
// elsewhere (and on the cells beyond an odd edge)
template <bool RESIDUAL>
__global__ __launch_bounds__(256) void k_res_bwd(Img m0, Img m1, const __bf16* __restrict__ XS,
                                                 const float* __restrict__ re, const double* __restrict__ sums,
                                                 float gate, float* __restrict__ G4) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)m1.H * m1.W) return;
  const int y = (int)(idx / m1.W), x = (int)(idx % m1.W);
  const double b = sums[1];
  const float scale = b > 0.0 ? (float)(2.0 / b) : 0.f;
  float g[NET_HO];
#pragma unroll
  for (int s = 0; s < NET_HO; ++s) {
    const int fy = 2 * y + (s >> 1), fx = 2 * x + (s & 1);
    g[s] = 0.f;
    if (fy < m0.H && fx < m0.W && active(m1, XS, fy, fx, gate))
      g[s] = scale * (RESIDUAL ? apply_A(m0, m1, XS, fy, fx,
                                         [&](int yy, int xx) { return re[(int64_t)yy * m0.W + xx]; })
                               : re[(int64_t)fy * m0.W + fx]);
  }
  *reinterpret_cast<float4*>(G4 + img_pix(m1, y, x) * NET_HO) = make_float4(g[0], g[1], g[2], g[3]);
}

// Head forward for inference on MFMA: the 4-output 3x3 convolution as the
// implicit GEMM of k_conv3x3 with one 16-row M-tile whose rows 0-3 hold the
// head weights (rows 4-15 zero), the same strip / LDS-ring structure, and the
// bias and activity gate applied to rows 0-3.  Lanes 0-15 of each wave hold
// rows 0-3 of their 16 pixels.  (Training keeps k_head_fwd, which also forms
// the loss gradient; the two differ only in the FP32 summation order.)
__global__ __launch_bounds__(256) void k_head_mfma(Img m0, Img m1, const __bf16* __restrict__ U1,
                                                   const __bf16* __restrict__ wh, const float* __restrict__ params,
                                                   const __bf16* __restrict__ XS, float gate,
                                                   float* __restrict__ delta) {
  constexpr int KB = 9, RX = 66, NCH = RX * CG, NLD = (NCH + 255) / 256;
  __shared__ bf16x8 sRow[3][NCH];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int x0 = blockIdx.x * 64, y0 = blockIdx.y * CONV_RB;
  const int y1 = min(y0 + CONV_RB, m1.H);
  bf16x8 a[KB];
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) a[kb] = lr < NET_HO ? load8(wh + (lr * 9 + kb) * NET_C + 8 * lg) : zero8();
  float bh[NET_HO];
#pragma unroll
  for (int s = 0; s < NET_HO; ++s) bh[s] = params[NP_BH + s];

  bf16x8 pre[NLD];
  auto fetch = [&](int yi) {
#pragma unroll
    for (int q = 0; q < NLD; ++q) {
      const int t = tid + q * 256;
      pre[q] = zero8();
      if (t < NCH) {
        const int px = t / CG, c = t % CG;
        const int xi = x0 - 1 + px;
        if (yi <= m1.H && xi <= m1.W) pre[q] = load8(U1 + img_pix(m1, yi, xi) * NET_C + 8 * c);
      }
    }
  };
  auto store = [&](int yi) {
    bf16x8* R = sRow[(yi + 3) % 3];
#pragma unroll
    for (int q = 0; q < NLD; ++q) {
      const int t = tid + q * 256;
      if (t < NCH) R[t] = pre[q];
    }
  };
  fetch(y0 - 1);
  store(y0 - 1);
  fetch(y0);
  store(y0);
  fetch(y0 + 1);
  const int xo = 16 * w + lr;
  for (int y = y0; y < y1; ++y) {
    store(y + 1);
    __syncthreads();
    if (y + 1 < y1) fetch(y + 2);
    f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      const int dy = kb / 3 - 1, dx = kb % 3 - 1;
      acc = mfma16(a[kb], sRow[(y + dy + 3) % 3][(xo + 1 + dx) * CG + lg], acc);
    }
    const int x = x0 + xo;
    if (lg == 0 && x < m1.W) {
#pragma unroll
      for (int s = 0; s < NET_HO; ++s) {
        const int fy = 2 * y + (s >> 1), fx = 2 * x + (s & 1);
        if (fy < m0.H && fx < m0.W)
          delta[(int64_t)fy * m0.W + fx] = active(m1, XS, fy, fx, gate) ? acc[s] + bh[s] : 0.f;
      }
    }
    __syncthreads();
  }
}

// GU1[p][ci] = (sum_tap sum_s G4[p - off(tap)][s] Wh[s][tap][ci]) * (U1[p][ci] > 0)
__global__ __launch_bounds__(256) void k_head_bwd(Img m1, const float* __restrict__ G4, const __bf16* __restrict__ wh,
                                                  const __bf16* __restrict__ U1, __bf16* __restrict__ GU1) {
  __shared__ float sw[NET_HO * 9 * NET_C];
  for (int t = threadIdx.x; t < NET_HO * 9 * NET_C; t += 256) sw[t] = bf(wh[t]);
  __syncthreads();
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)m1.H * m1.W) return;
  const int y = (int)(idx / m1.W), x = (int)(idx % m1.W);
  float g[9][NET_HO];
#pragma unroll
  for (int tap = 0; tap < 9; ++tap) {
    const float4 v =
        *reinterpret_cast<const float4*>(G4 + img_pix(m1, y - (tap / 3 - 1), x - (tap % 3 - 1)) * NET_HO);
    g[tap][0] = v.x;
    g[tap][1] = v.y;
    g[tap][2] = v.z;
    g[tap][3] = v.w;
  }
  const int64_t p = img_pix(m1, y, x);
#pragma unroll
  for (int q = 0; q < CG; ++q) {
    const bf16x8 um = load8(U1 + p * NET_C + 8 * q);
    bf16x8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      float s = 0.f;
#pragma unroll
      for (int tap = 0; tap < 9; ++tap)
#pragma unroll
        for (int h = 0; h < NET_HO; ++h) s += g[tap][h] * sw[(h * 9 + tap) * NET_C + 8 * q + j];
      o[j] = (__bf16)(bf(um[j]) > 0.f ? s : 0.f);
    }
    store8(GU1 + p * NET_C + 8 * q, o);
  }
}

constexpr int WG_RB = 4;       // rows per LDS stage
constexpr int WG_ROWS = 32;    // rows per block
constexpr int WG_GS = NET_C + 2;  // LDS stride of a G pixel (bank spread)

constexpr int CIN = NET_C;
constexpr int wg_ntiles() { return (9 * CIN + 1 + 15) / 16; }

__global__ __launch_bounds__(256) void k_conv_wgrad(Img m, const __bf16* __restrict__ G, const __bf16* __restrict__ X,
                                                    float* __restrict__ part) {
  constexpr int NT = wg_ntiles();
  constexpr int NTW = (NT + 3) / 4;  // N-tiles per wave
  constexpr int XS = CIN + 2;
  __shared__ __bf16 sG[WG_RB * 64 * WG_GS];
  __shared__ __bf16 sX[(WG_RB + 2) * 66 * XS];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6, lr = lane & 15, lg = lane >> 4;
  const int x0 = blockIdx.x * 64;
  const int y0 = blockIdx.y * WG_ROWS;
  f32x4 acc[NTW][2];
#pragma unroll
  for (int t = 0; t < NTW; ++t) acc[t][0] = acc[t][1] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int yend = min(y0 + WG_ROWS, m.H);
  for (int ys = y0; ys < yend; ys += WG_RB) {
    __syncthreads();
    // G tile: rows ys..ys+3, pixels x0..x0+63 (zero outside the image)
    for (int e = tid; e < WG_RB * 64 * (NET_C / 8); e += 256) {
      const int q = e % (NET_C / 8), px = (e / (NET_C / 8)) % 64, r = e / (64 * (NET_C / 8));
      const int y = ys + r, x = x0 + px;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (y < m.H && x < m.W) v = *reinterpret_cast<const uint4*>(G + img_pix(m, y, x) * NET_C + 8 * q);
      unsigned* d = reinterpret_cast<unsigned*>(sG + (r * 64 + px) * WG_GS + 8 * q);
      d[0] = v.x;
      d[1] = v.y;
      d[2] = v.z;
      d[3] = v.w;
    }
    // X tile: rows ys-1..ys+4, pixels x0-1..x0+64 of the bordered image
    for (int e = tid; e < (WG_RB + 2) * 66 * (CIN / 8); e += 256) {
      const int q = e % (CIN / 8), px = (e / (CIN / 8)) % 66, r = e / (66 * (CIN / 8));
      const int y = ys - 1 + r, x = x0 - 1 + px;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (y <= m.H && x <= m.W) v = *reinterpret_cast<const uint4*>(X + img_pix(m, y, x) * CIN + 8 * q);
      unsigned* d = reinterpret_cast<unsigned*>(sX + (r * 66 + px) * XS + 8 * q);
      d[0] = v.x;
      d[1] = v.y;
      d[2] = v.z;
      d[3] = v.w;
    }
    __syncthreads();
    const int nkb = 2 * min(WG_RB, m.H - ys);
    for (int kb = 0; kb < nkb; ++kb) {
      const int r = kb >> 1, pxb = 32 * (kb & 1) + 8 * lg;
      bf16x8 a0, a1;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        a0[j] = sG[(r * 64 + pxb + j) * WG_GS + lr];
        a1[j] = sG[(r * 64 + pxb + j) * WG_GS + 16 + lr];
      }
#pragma unroll
      for (int t = 0; t < NTW; ++t) {
        const int n = w + 4 * t;
        if (n < NT) {
          const int k = 16 * n + lr;
          bf16x8 b;
          if (k < 9 * CIN) {
            const int tap = k / CIN, ci = k % CIN;
            const int dy = tap / 3 - 1, dx = tap % 3 - 1;
#pragma unroll
            for (int j = 0; j < 8; ++j) b[j] = sX[((r + 1 + dy) * 66 + (pxb + j + 1 + dx)) * XS + ci];
          } else {
            const __bf16 one = (__bf16)(k == 9 * CIN ? 1.f : 0.f);
#pragma unroll
            for (int j = 0; j < 8; ++j) b[j] = one;
          }
          acc[t][0] = mfma16(a0, b, acc[t][0]);
          acc[t][1] = mfma16(a1, b, acc[t][1]);
        }
      }
    }
  }
  const int blk = blockIdx.y * gridDim.x + blockIdx.x;
#pragma unroll
  for (int t = 0; t < NTW; ++t) {
    const int n = w + 4 * t;
    if (n >= NT) continue;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr)
        part[((int64_t)blk * NET_C + 16 * mt + 4 * lg + rr) * (NT * 16) + 16 * n + lr] = acc[t][mt][rr];
  }
}

constexpr int HW_PPB = 4096;  // level-1 pixels per block of the head weight gradient
constexpr int HW_N = 9 * NET_C + 1;  // (tap, ci) and the bias, per head output

// head weight gradient partials: thread o < 289 of a block sums, over its
// pixels, G4[p][s] U1[p + off(tap)][ci] (o = 32 tap + ci) or G4[p][s] (o = 288)
// for all four outputs s; part[block][s][o]
__global__ __launch_bounds__(256) void k_head_wgrad(Img m1, const float* __restrict__ G4, const __bf16* __restrict__ U1,
                                                    float* __restrict__ part) {
  const int64_t npx = (int64_t)m1.H * m1.W;
  const int64_t p0 = (int64_t)blockIdx.x * HW_PPB, p1 = min(p0 + HW_PPB, npx);
  for (int o = threadIdx.x; o < HW_N; o += 256) {
    float s[NET_HO] = {0.f, 0.f, 0.f, 0.f};
    const int tap = o / NET_C, ci = o % NET_C;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    for (int64_t q = p0; q < p1; ++q) {
      const int y = (int)(q / m1.W), x = (int)(q % m1.W);
      const float4 g = *reinterpret_cast<const float4*>(G4 + img_pix(m1, y, x) * NET_HO);
      const float xv = o < 9 * NET_C ? bf(U1[img_pix(m1, y + dy, x + dx) * NET_C + ci]) : 1.f;
      s[0] += g.x * xv;
      s[1] += g.y * xv;
      s[2] += g.z * xv;
      s[3] += g.w * xv;
    }
#pragma unroll
    for (int h = 0; h < NET_HO; ++h) part[((int64_t)blockIdx.x * NET_HO + h) * HW_N + o] = s[h];
  }
}

__global__ void k_reduce_conv_grad(int nblocks, const float* __restrict__ part, float* __restrict__ gW,
                                   float* __restrict__ gB) {
  constexpr int NPAD = wg_ntiles() * 16, NK = 9 * NET_C;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= NET_C * (NK + 1)) return;
  const int co = j / (NK + 1), n = j % (NK + 1);
  float s = 0.f;
  for (int b = 0; b < nblocks; ++b) s += part[((int64_t)b * NET_C + co) * NPAD + n];
  if (n < NK)
    gW[co * NK + n] = s;
  else
    gB[co] = s;
}

__global__ void k_reduce_head_grad(int nblocks, const float* __restrict__ part, float* __restrict__ grad) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= NET_HO * HW_N) return;
  const int h = j / HW_N, o = j % HW_N;
  float s = 0.f;
  for (int b = 0; b < nblocks; ++b) s += part[(int64_t)b * NET_HO * HW_N + j];
  if (o < 9 * NET_C)
    grad[NP_WH + h * 9 * NET_C + o] = s;
  else
    grad[NP_BH + h] = s;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = isfinite(g[i]) ? g[i] : 0.f;
  const float m = b1 * m1[i] + (1.f - b1) * gi;
  const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = m;
  m2[i] = v;
  p[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

__global__ void k_pack_weights(const float* __restrict__ P, NetWeights w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < NET_C * 9 * NET_C) {
    // flipped: out channel ci, tap t', in channel co  <-  W[co][8 - t'][ci]
    const int ci = i / (9 * NET_C), tp = (i / NET_C) % 9, co = i % NET_C;
    const int src = co * 9 * NET_C + (8 - tp) * NET_C + ci;
#pragma unroll
    for (int l = 1; l <= NET_CONVS; ++l) {
      w.wp[l][i] = (__bf16)P[np_w(l) + i];
      if (l >= 2) w.wf[l][i] = (__bf16)P[np_w(l) + src];
    }
  }
  if (i < NET_HO * 9 * NET_C) w.wh[i] = (__bf16)P[NP_WH + i];
}

__global__ void k_fill_bf16(__bf16* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (__bf16)0.f;
}

__device__ __forceinline__ float finite_or_zero(double v) { return isfinite(v) ? (float)v : 0.f; }

// The 8 features of a domain cell: T^n / max|T^n|; the previous increment
// rescaled to this dt, / E; J / J_ref - 1; log(1 + dt abar_xi g11) and
// log(1 + dt abar_eta g22) (the face diffusion numbers); dt r(T^n) / E; 1;
// the largest |feature 5| over the 3x3 neighbourhood.  Stored in the
// level-1 block layout (cell
// (y, x) is channel group 2 (y & 1) + (x & 1) of level-1 pixel (y/2, x/2))
__global__ void k_features(FeatIn f, Img m0, Img m1, __bf16* __restrict__ XS) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)m0.H * m0.W) return;
  const int y = (int)(idx / m0.W), x = (int)(idx % m0.W);
  const Geo& g = f.g;
  const int i = g.lo[0] + x, j = g.lo[1] + y;
  const int64_t c = at(g, i, j, 0);
  const double ts = f.scales[0], es = f.scales[1];
  const double it = ts > 0.0 ? 1.0 / ts : 0.0, ie = es > 0.0 ? 1.0 / es : 0.0;
  const double u = f.u[c];
  bf16x8 o;
  o[0] = (__bf16)finite_or_zero(u * it);
  o[1] = (__bf16)finite_or_zero((u - f.um1[c]) * ie * f.dt_ratio);
  o[2] = (__bf16)finite_or_zero(f.jac[c] / f.j_ref - 1.0);
  const double ax = face_mean<double>(f.harmonic, f.a[c], f.a[c + 1]);
  const double ay = face_mean<double>(f.harmonic, f.a[c], f.a[c + g.pitch]);
  o[3] = (__bf16)finite_or_zero(log1p(fmax(0.0, f.dt * f.g11[c] * ax)));
  o[4] = (__bf16)finite_or_zero(log1p(fmax(0.0, f.dt * f.g22[c] * ay)));
  o[5] = (__bf16)finite_or_zero(f.dt * (double)f.r32[c] * ie);
  o[6] = (__bf16)1.f;
  // feature 7: the largest |feature 5| over the 3x3 neighbourhood (cells
  // outside the domain count 0), each rounded as feature 5 is
  float act = 0.f;
  for (int tap = 0; tap < 9; ++tap) {
    const int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
    if (yy >= 0 && yy < m0.H && xx >= 0 && xx < m0.W) {
      const __bf16 v = (__bf16)finite_or_zero(f.dt * (double)f.r32[at(g, g.lo[0] + xx, g.lo[1] + yy, 0)] * ie);
      act = fmaxf(act, fabsf(bf(v)));
    }
  }
  o[7] = (__bf16)act;
  store8(XS + img_pix(m1, y >> 1, x >> 1) * NET_C + 8 * (2 * (y & 1) + (x & 1)), o);
}

__global__ void k_target(Geo g, Img m, const double* __restrict__ unew, const double* __restrict__ uold,
                         const double* __restrict__ scales, float* __restrict__ tgt) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)m.H * m.W) return;
  const int y = (int)(idx / m.W), x = (int)(idx % m.W);
  const int64_t c = at(g, g.lo[0] + x, g.lo[1] + y, 0);
  const double es = scales[1];
  tgt[idx] = finite_or_zero((unew[c] - uold[c]) * (es > 0.0 ? 1.0 / es : 0.0));
}

__global__ void k_apply_prediction(Geo g, const double* __restrict__ u, const float* __restrict__ delta,
                                   const double* __restrict__ scales, double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ntot = (int64_t)g.n[0] * g.n[1];
  if (e >= ntot) return;
  const int i = (int)(e % g.n[0]), j = (int)(e / g.n[0]);
  const int64_t c = at(g, i, j, 0);
  double v = u[c];
  if (i >= g.lo[0] && i <= g.hi[0] && j >= g.lo[1] && j <= g.hi[1]) {
    const int W = g.hi[0] - g.lo[0] + 1;
    v += scales[1] * (double)delta[(int64_t)(j - g.lo[1]) * W + (i - g.lo[0])];
  }
  out[c] = v;
}

template <typename T>
__global__ void k_maxabs_partial_t(Geo g, const T* __restrict__ u, double mul, double* __restrict__ part) {
  __shared__ double s[256];
  const int ni = g.hi[0] - g.lo[0] + 1;
  const int64_t n = (int64_t)ni * (g.hi[1] - g.lo[1] + 1);
  double mx = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    const double v = fabs((double)u[at(g, g.lo[0] + (int)(e % ni), g.lo[1] + (int)(e / ni), 0)]);
    if (v > mx) mx = v;  // NaN does not raise the maximum
  }
  s[threadIdx.x] = mx * mul;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void k_max_final(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double s[256];
  double mx = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) mx = fmax(mx, part[i]);
  s[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

dim3 grid1(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

dim3 conv_grid(Img m) { return dim3((m.W + 63) / 64, (m.H + CONV_RB - 1) / CONV_RB); }

}  // namespace

void launch_conv3x3(int mode, Img m, const __bf16* X, const __bf16* Wp, const float* bias, const __bf16* mask,
                    __bf16* Y, hipStream_t s) {
  if (mode == 0)
    hipLaunchKernelGGL(k_conv3x3<0>, conv_grid(m), dim3(256), 0, s, m, X, Wp, bias, mask, Y);
  else if (mode == 1)
    hipLaunchKernelGGL(k_conv3x3<1>, conv_grid(m), dim3(256), 0, s, m, X, Wp, bias, mask, Y);
  else
    hipLaunchKernelGGL(k_conv3x3<2>, conv_grid(m), dim3(256), 0, s, m, X, Wp, bias, mask, Y);
}

void launch_pool(Img m, const __bf16* X, __bf16* P, hipStream_t s) {
  const Img mc = img_half(m);
  hipLaunchKernelGGL(k_pool, grid1((int64_t)mc.H * mc.W * CG), dim3(256), 0, s, m, mc, X, P);
}

void launch_up_add(Img m, const __bf16* C, const __bf16* S, __bf16* A, hipStream_t s) {
  hipLaunchKernelGGL(k_up_add, grid1((int64_t)m.H * m.W * CG), dim3(256), 0, s, m, img_half(m), C, S, A);
}

void launch_grad_skip(Img m, const __bf16* GA, const __bf16* GP, const __bf16* E, __bf16* GE, hipStream_t s) {
  hipLaunchKernelGGL(k_grad_skip, grid1((int64_t)m.H * m.W * CG), dim3(256), 0, s, m, img_half(m), GA, GP, E, GE);
}

void launch_grad_up(Img m, const __bf16* GA, const __bf16* U, __bf16* GU, hipStream_t s) {
  const Img mc = img_half(m);
  hipLaunchKernelGGL(k_grad_up, grid1((int64_t)mc.H * mc.W * CG), dim3(256), 0, s, m, mc, GA, U, GU);
}

void launch_s2d(Img m0, const __bf16* X0, __bf16* XS, hipStream_t s) {
  hipLaunchKernelGGL(k_s2d, grid1((int64_t)m0.H * m0.W), dim3(256), 0, s, m0, img_half(m0), X0, XS);
}

void launch_head_infer(Img m0, const __bf16* U1, const __bf16* wh, const float* params, const __bf16* XS, float gate,
                       float* delta, hipStream_t s) {
  const Img m1 = img_half(m0);
  hipLaunchKernelGGL(k_head_mfma, conv_grid(m1), dim3(256), 0, s, m0, m1, U1, wh, params, XS, gate, delta);
}

void launch_head_fwd(Img m0, const __bf16* U1, const __bf16* wh, const float* params, const __bf16* XS, float gate,
                     float* delta, hipStream_t s) {
  const Img m1 = img_half(m0);
  hipLaunchKernelGGL(k_head_fwd, grid1((int64_t)m1.H * m1.W), dim3(256), 0, s, m0, m1, U1, wh, params, XS, gate,
                     delta);
}

int loss_parts(Img m0) { return (int)(((int64_t)m0.H * m0.W + 255) / 256); }

void launch_loss(int kind, Img m0, const __bf16* XS, const float* delta, const float* target, float gate, float* re,
                 double* part, double* sums, double* loss, float* G4, hipStream_t s) {
  const Img m1 = img_half(m0);
  const int np = loss_parts(m0);
  if (kind == 1)
    hipLaunchKernelGGL(k_res_fwd<true>, dim3(np), dim3(256), 0, s, m0, m1, XS, delta, target, re, part);
  else
    hipLaunchKernelGGL(k_res_fwd<false>, dim3(np), dim3(256), 0, s, m0, m1, XS, delta, target, re, part);
  hipLaunchKernelGGL(k_res_sums, dim3(1), dim3(256), 0, s, (const double*)part, np, sums, loss);
  if (kind == 1)
    hipLaunchKernelGGL(k_res_bwd<true>, grid1((int64_t)m1.H * m1.W), dim3(256), 0, s, m0, m1, XS, (const float*)re,
                       (const double*)sums, gate, G4);
  else
    hipLaunchKernelGGL(k_res_bwd<false>, grid1((int64_t)m1.H * m1.W), dim3(256), 0, s, m0, m1, XS,
                       (const float*)re, (const double*)sums, gate, G4);
}

void launch_head_bwd(Img m0, const float* G4, const __bf16* wh, const __bf16* U1, __bf16* GU1, hipStream_t s) {
  const Img m1 = img_half(m0);
  hipLaunchKernelGGL(k_head_bwd, grid1((int64_t)m1.H * m1.W), dim3(256), 0, s, m1, G4, wh, U1, GU1);
}

int conv_wgrad_npad() { return wg_ntiles() * 16; }

int conv_wgrad_blocks(Img m) { return ((m.W + 63) / 64) * ((m.H + WG_ROWS - 1) / WG_ROWS); }

int launch_conv_wgrad(Img m, const __bf16* G, const __bf16* X, float* part, hipStream_t s) {
  const dim3 grid((m.W + 63) / 64, (m.H + WG_ROWS - 1) / WG_ROWS);
  hipLaunchKernelGGL(k_conv_wgrad, grid, dim3(256), 0, s, m, G, X, part);
  return (int)(grid.x * grid.y);
}

int head_wgrad_blocks(Img m1) { return (int)(((int64_t)m1.H * m1.W + HW_PPB - 1) / HW_PPB); }

int launch_head_wgrad(Img m1, const float* G4, const __bf16* U1, float* part, hipStream_t s) {
  const int nb = head_wgrad_blocks(m1);
  hipLaunchKernelGGL(k_head_wgrad, dim3(nb), dim3(256), 0, s, m1, G4, U1, part);
  return nb;
}

void launch_reduce_conv_grad(int nblocks, const float* part, float* gW, float* gB, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_conv_grad, grid1(NET_C * (9 * NET_C + 1)), dim3(256), 0, s, nblocks, part, gW, gB);
}

void launch_reduce_head_grad(int nblocks, const float* part, float* grad, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_head_grad, grid1(NET_HO * HW_N), dim3(256), 0, s, nblocks, part, grad);
}

void launch_adam(float* params, const float* grad, float* m1, float* m2, int n, float lr, float b1, float b2,
                 float eps, int step, hipStream_t s) {
  const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
  hipLaunchKernelGGL(k_adam, grid1(n), dim3(256), 0, s, params, grad, m1, m2, n, lr, b1, b2, eps, c1, c2);
}

void launch_pack_weights(const float* params, NetWeights w, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_weights, grid1(NET_C * 9 * NET_C), dim3(256), 0, s, params, w);
}

void launch_fill_bf16(__bf16* p, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_bf16, grid1(n), dim3(256), 0, s, p, n);
}

void launch_features(const FeatIn& f, __bf16* XS, hipStream_t s) {
  const Img m0{f.g.hi[1] - f.g.lo[1] + 1, f.g.hi[0] - f.g.lo[0] + 1};
  hipLaunchKernelGGL(k_features, grid1((int64_t)m0.H * m0.W), dim3(256), 0, s, f, m0, img_half(m0), XS);
}

void launch_target(Geo g, const double* unew, const double* uold, const double* scales, float* tgt, hipStream_t s) {
  const Img m{g.hi[1] - g.lo[1] + 1, g.hi[0] - g.lo[0] + 1};
  hipLaunchKernelGGL(k_target, grid1((int64_t)m.H * m.W), dim3(256), 0, s, g, m, unew, uold, scales, tgt);
}

void launch_apply_prediction(Geo g, const double* u, const float* delta, const double* scales, double* out,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_apply_prediction, grid1((int64_t)g.n[0] * g.n[1]), dim3(256), 0, s, g, u, delta, scales, out);
}

void launch_domain_maxabs(Geo g, const double* u, double* part, double* out, hipStream_t s) {
  hipLaunchKernelGGL((k_maxabs_partial_t<double>), dim3(1024), dim3(256), 0, s, g, u, 1.0, part);
  hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, s, (const double*)part, 1024, out);
}

void launch_domain_maxabs32(Geo g, const float* r, double mul, double* part, double* out, hipStream_t s) {
  hipLaunchKernelGGL((k_maxabs_partial_t<float>), dim3(1024), dim3(256), 0, s, g, r, mul, part);
  hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, s, (const double*)part, 1024, out);
}

}  // namespace cdiff
