// learn.hip -- kernels of the learned warm start (see learn.h).
//
// The three 32-channel layers run on MFMA (v_mfma_f32_16x16x32_bf16, BF16
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
// The 1-channel head (288 MACs per pixel) runs on MFMA for inference (one
// M-tile, row 0 real) and on the vector ALUs in training, where it also forms
// the loss gradient.
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

__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u)); }

__device__ __forceinline__ float bf(const __bf16 v) { return (float)v; }

constexpr int CONV_RB = 16;  // output rows per forward block

// 3x3 convolution, CIN -> 32 channels, as an implicit GEMM on MFMA.  A block
// of 4 waves owns a 64-pixel-wide strip of CONV_RB rows; wave w computes the
// 16-pixel N-tile w of every row for both 16-channel M-tiles, with the
// weights (A operands) in registers for the whole strip.  The input rows pass
// through a 3-row LDS ring (66 pixels x CIN channels): each input row is read
// from memory once per strip, the next row's load is in flight while the
// current row is computed, and the nine taps read their B operands (8
// channels of one pixel per lane) from LDS.  The MFMA chain per output tile
// is the same as a direct evaluation (taps in order), so results do not
// depend on the staging.
//   MODE 0: y = relu(acc + bias)      MODE 1: y = acc where mask > 0, else 0
template <int CIN, int MODE>
__global__ __launch_bounds__(256) void k_conv3x3(Img m, const __bf16* __restrict__ X, const __bf16* __restrict__ Wp,
                                                 const float* __restrict__ bias, const __bf16* __restrict__ mask,
                                                 __bf16* __restrict__ Y) {
  constexpr int KB = CIN == 32 ? 9 : 3;  // K-blocks of 32
  constexpr int RX = 66;                 // ring row: pixels x0-1 .. x0+64
  constexpr int CH = CIN / 8;            // 16-byte chunks per pixel
  constexpr int NCH = RX * CH;           // chunks per ring row
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
        const int px = t / CH, c = t % CH;
        const int xi = x0 - 1 + px;
        if (yi <= m.H && xi <= m.W) pre[q] = load8(X + img_pix(m, yi, xi) * CIN + 8 * c);
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
      const int tap = CIN == 32 ? kb : 4 * kb + lg;
      const int c0 = CIN == 32 ? lg : 0;
      const int dy = tap / 3 - 1, dx = tap % 3 - 1;
      const bf16x8 bv = tap < 9 ? sRow[(y + dy + 3) % 3][(xo + 1 + dx) * CH + c0] : zero8();
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
          else
            v = bf(mask[p * NET_C + co + r]) > 0.f ? v : 0.f;
          o[r] = (__bf16)v;
        }
        *reinterpret_cast<bf16x4*>(Y + p * NET_C + co) = o;
      }
    }
    __syncthreads();  // row y-1's slot is reloaded next
  }
}

// A pixel is active when the explicit-Euler increment (feature 5, scaled by
// its domain maximum) exceeds `gate` in its 3x3 neighbourhood; elsewhere the
// prediction is 0 (no change), so the network adds no background error.
__device__ __forceinline__ bool active(const Img& m, const __bf16* __restrict__ X0, int y, int x, float gate) {
  float e = 0.f;
  for (int tap = 0; tap < 9; ++tap)
    e = fmaxf(e, fabsf(bf(X0[img_pix(m, y + tap / 3 - 1, x + tap % 3 - 1) * NET_C0 + 5])));
  return e > gate;
}

// head forward (+ loss gradient when a target is given)
__global__ __launch_bounds__(256) void k_head_fwd(Img m, const __bf16* __restrict__ X3, const __bf16* __restrict__ w4,
                                                  const float* __restrict__ params, const __bf16* __restrict__ X0,
                                                  float gate, float* __restrict__ delta,
                                                  const float* __restrict__ target, float* __restrict__ G4,
                                                  double* __restrict__ loss_part) {
  __shared__ float sw[9 * NET_C];
  __shared__ double sred[256 / WAVE];
  for (int t = threadIdx.x; t < 9 * NET_C; t += 256) sw[t] = bf(w4[t]);
  __syncthreads();
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t npx = (int64_t)m.H * m.W;
  double l = 0.0;
  if (idx < npx) {
    const int y = (int)(idx / m.W), x = (int)(idx % m.W);
    float acc = params[NP_B4];
    for (int tap = 0; tap < 9; ++tap) {
      const __bf16* src = X3 + img_pix(m, y + tap / 3 - 1, x + tap % 3 - 1) * NET_C;
#pragma unroll
      for (int q = 0; q < NET_C / 8; ++q) {
        const bf16x8 v = load8(src + 8 * q);
#pragma unroll
        for (int j = 0; j < 8; ++j) acc += bf(v[j]) * sw[tap * NET_C + 8 * q + j];
      }
    }
    const bool on = active(m, X0, y, x, gate);
    if (!on) acc = 0.f;
    delta[idx] = acc;
    if (target) {
      const float d = acc - target[idx];
      G4[img_pix(m, y, x)] = on ? 2.f * d / (float)npx : 0.f;
      l = (double)d * (double)d;
    }
  }
  if (!target) return;
  l = wave_sum(l);
  if ((threadIdx.x & (WAVE - 1)) == 0) sred[threadIdx.x / WAVE] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < 256 / WAVE; ++i) s += sred[i];
    loss_part[blockIdx.x] = s;
  }
}

// Head forward for inference on MFMA: the 1-channel 3x3 convolution as the
// implicit GEMM of k_conv3x3 with one 16-row M-tile whose row 0 holds the head
// weights (rows 1-15 zero), the same strip / LDS-ring structure, and the bias
// and activity gate applied to row 0.  Lanes 0-15 of each wave hold row 0 of
// their 16 pixels.  (Training keeps k_head_fwd, which also forms the loss
// gradient; the two differ only in the FP32 summation order.)
__global__ __launch_bounds__(256) void k_head_mfma(Img m, const __bf16* __restrict__ X3, const __bf16* __restrict__ w4,
                                                   const float* __restrict__ params, const __bf16* __restrict__ X0,
                                                   float gate, float* __restrict__ delta) {
  constexpr int KB = 9, RX = 66, CH = NET_C / 8, NCH = RX * CH, NLD = (NCH + 255) / 256;
  __shared__ bf16x8 sRow[3][NCH];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int x0 = blockIdx.x * 64, y0 = blockIdx.y * CONV_RB;
  const int y1 = min(y0 + CONV_RB, m.H);
  bf16x8 a[KB];
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) a[kb] = lr == 0 ? load8(w4 + kb * NET_C + 8 * lg) : zero8();
  const float b4 = params[NP_B4];

  bf16x8 pre[NLD];
  auto fetch = [&](int yi) {
#pragma unroll
    for (int q = 0; q < NLD; ++q) {
      const int t = tid + q * 256;
      pre[q] = zero8();
      if (t < NCH) {
        const int px = t / CH, c = t % CH;
        const int xi = x0 - 1 + px;
        if (yi <= m.H && xi <= m.W) pre[q] = load8(X3 + img_pix(m, yi, xi) * NET_C + 8 * c);
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
      acc = mfma16(a[kb], sRow[(y + dy + 3) % 3][(xo + 1 + dx) * CH + lg], acc);
    }
    const int x = x0 + xo;
    if (lg == 0 && x < m.W) {
      const float v = acc[0] + b4;
      delta[(int64_t)y * m.W + x] = active(m, X0, y, x, gate) ? v : 0.f;
    }
    __syncthreads();
  }
}

// G3pre[p][ci] = (sum_tap G4[p - off(tap)] W4[tap][ci]) * (X3[p][ci] > 0)
__global__ __launch_bounds__(256) void k_head_bwd(Img m, const float* __restrict__ G4, const __bf16* __restrict__ w4,
                                                  const __bf16* __restrict__ X3, __bf16* __restrict__ G3) {
  __shared__ float sw[9 * NET_C];
  for (int t = threadIdx.x; t < 9 * NET_C; t += 256) sw[t] = bf(w4[t]);
  __syncthreads();
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)m.H * m.W) return;
  const int y = (int)(idx / m.W), x = (int)(idx % m.W);
  float g[9];
#pragma unroll
  for (int tap = 0; tap < 9; ++tap) g[tap] = G4[img_pix(m, y - (tap / 3 - 1), x - (tap % 3 - 1))];
  const int64_t p = img_pix(m, y, x);
#pragma unroll
  for (int q = 0; q < NET_C / 8; ++q) {
    const bf16x8 xm = load8(X3 + p * NET_C + 8 * q);
    bf16x8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      float s = 0.f;
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) s += g[tap] * sw[tap * NET_C + 8 * q + j];
      o[j] = (__bf16)(bf(xm[j]) > 0.f ? s : 0.f);
    }
    *reinterpret_cast<uint4*>(G3 + p * NET_C + 8 * q) = __builtin_bit_cast(uint4, o);
  }
}

constexpr int WG_RB = 4;       // rows per LDS stage
constexpr int WG_ROWS = 32;    // rows per block
constexpr int WG_GS = NET_C + 2;  // LDS stride of a G pixel (bank spread)

template <int CIN>
constexpr int wg_ntiles() {
  return (9 * CIN + 1 + 15) / 16;
}

template <int CIN>
__global__ __launch_bounds__(256) void k_conv_wgrad(Img m, const __bf16* __restrict__ G, const __bf16* __restrict__ X,
                                                    float* __restrict__ part) {
  constexpr int NT = wg_ntiles<CIN>();
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

constexpr int HW_PPB = 4096;  // pixels per block of the head weight gradient

__global__ __launch_bounds__(256) void k_head_wgrad(Img m, const float* __restrict__ G4, const __bf16* __restrict__ X3,
                                                    float* __restrict__ part) {
  const int64_t npx = (int64_t)m.H * m.W;
  const int64_t p0 = (int64_t)blockIdx.x * HW_PPB, p1 = min(p0 + HW_PPB, npx);
  for (int o = threadIdx.x; o < 9 * NET_C + 1; o += 256) {
    float s = 0.f;
    if (o < 9 * NET_C) {
      const int tap = o / NET_C, ci = o % NET_C;
      const int dy = tap / 3 - 1, dx = tap % 3 - 1;
      for (int64_t q = p0; q < p1; ++q) {
        const int y = (int)(q / m.W), x = (int)(q % m.W);
        s += G4[img_pix(m, y, x)] * bf(X3[img_pix(m, y + dy, x + dx) * NET_C + ci]);
      }
    } else {
      for (int64_t q = p0; q < p1; ++q) s += G4[img_pix(m, (int)(q / m.W), (int)(q % m.W))];
    }
    part[(int64_t)blockIdx.x * (9 * NET_C + 1) + o] = s;
  }
}

template <int CIN>
__global__ void k_reduce_conv_grad(int nblocks, const float* __restrict__ part, float* __restrict__ gW,
                                   float* __restrict__ gB) {
  constexpr int NPAD = wg_ntiles<CIN>() * 16, NK = 9 * CIN;
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
  if (j >= 9 * NET_C + 1) return;
  float s = 0.f;
  for (int b = 0; b < nblocks; ++b) s += part[(int64_t)b * (9 * NET_C + 1) + j];
  grad[NP_W4 + j] = s;
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
  // w1p: [32][96]
  if (i < NET_C * 96) {
    const int co = i / 96, k = i % 96;
    w.w1p[i] = (__bf16)(k < 9 * NET_C0 ? P[NP_W1 + co * 9 * NET_C0 + k] : 0.f);
  }
  if (i < NET_C * 9 * NET_C) {
    w.w2p[i] = (__bf16)P[NP_W2 + i];
    w.w3p[i] = (__bf16)P[NP_W3 + i];
    // flipped: out channel ci, tap t', in channel co  <-  W[co][8 - t'][ci]
    const int ci = i / (9 * NET_C), tp = (i / NET_C) % 9, co = i % NET_C;
    const int src = co * 9 * NET_C + (8 - tp) * NET_C + ci;
    w.w2f[i] = (__bf16)P[NP_W2 + src];
    w.w3f[i] = (__bf16)P[NP_W3 + src];
  }
  if (i < 9 * NET_C) w.w4[i] = (__bf16)P[NP_W4 + i];
}

__global__ void k_fill_bf16(__bf16* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (__bf16)0.f;
}

__global__ void k_sum_loss(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double s[256];
  double a = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) a += part[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

__device__ __forceinline__ float finite_or_zero(double v) { return isfinite(v) ? (float)v : 0.f; }

__global__ void k_features(FeatIn f, Img m, __bf16* __restrict__ X0) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)m.H * m.W) return;
  const int y = (int)(idx / m.W), x = (int)(idx % m.W);
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
  o[7] = (__bf16)0.f;
  *reinterpret_cast<uint4*>(X0 + img_pix(m, y, x) * NET_C0) = __builtin_bit_cast(uint4, o);
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

}  // namespace

void launch_conv3x3(int cin, int mode, Img m, const __bf16* X, const __bf16* Wp, const float* bias,
                    const __bf16* mask, __bf16* Y, hipStream_t s) {
  const dim3 grid((m.W + 63) / 64, (m.H + CONV_RB - 1) / CONV_RB);
  if (cin == NET_C0)
    hipLaunchKernelGGL((k_conv3x3<NET_C0, 0>), grid, dim3(256), 0, s, m, X, Wp, bias, mask, Y);
  else if (mode == 0)
    hipLaunchKernelGGL((k_conv3x3<NET_C, 0>), grid, dim3(256), 0, s, m, X, Wp, bias, mask, Y);
  else
    hipLaunchKernelGGL((k_conv3x3<NET_C, 1>), grid, dim3(256), 0, s, m, X, Wp, bias, mask, Y);
}

void launch_head_infer(Img m, const __bf16* X3, const __bf16* w4, const float* params, const __bf16* X0, float gate,
                       float* delta, hipStream_t s) {
  const dim3 grid((m.W + 63) / 64, (m.H + CONV_RB - 1) / CONV_RB);
  hipLaunchKernelGGL(k_head_mfma, grid, dim3(256), 0, s, m, X3, w4, params, X0, gate, delta);
}
void launch_head_fwd(Img m, const __bf16* X3, const __bf16* w4, const float* params, const __bf16* X0, float gate,
                     float* delta, const float* target, float* G4, double* loss_part, int* nparts, hipStream_t s) {
  const int64_t npx = (int64_t)m.H * m.W;
  hipLaunchKernelGGL(k_head_fwd, grid1(npx), dim3(256), 0, s, m, X3, w4, params, X0, gate, delta, target, G4,
                     loss_part);
  if (nparts) *nparts = (int)((npx + 255) / 256);
}

void launch_head_bwd(Img m, const float* G4, const __bf16* w4, const __bf16* X3, __bf16* G3, hipStream_t s) {
  hipLaunchKernelGGL(k_head_bwd, grid1((int64_t)m.H * m.W), dim3(256), 0, s, m, G4, w4, X3, G3);
}

int conv_wgrad_npad(int cin) { return cin == NET_C0 ? wg_ntiles<NET_C0>() * 16 : wg_ntiles<NET_C>() * 16; }

int launch_conv_wgrad(int cin, Img m, const __bf16* G, const __bf16* X, float* part, hipStream_t s) {
  const dim3 grid((m.W + 63) / 64, (m.H + WG_ROWS - 1) / WG_ROWS);
  if (cin == NET_C0)
    hipLaunchKernelGGL((k_conv_wgrad<NET_C0>), grid, dim3(256), 0, s, m, G, X, part);
  else
    hipLaunchKernelGGL((k_conv_wgrad<NET_C>), grid, dim3(256), 0, s, m, G, X, part);
  return (int)(grid.x * grid.y);
}

int launch_head_wgrad(Img m, const float* G4, const __bf16* X3, float* part, hipStream_t s) {
  const int nb = (int)(((int64_t)m.H * m.W + HW_PPB - 1) / HW_PPB);
  hipLaunchKernelGGL(k_head_wgrad, dim3(nb), dim3(256), 0, s, m, G4, X3, part);
  return nb;
}

void launch_reduce_conv_grad(int cin, int nblocks, const float* part, float* gW, float* gB, hipStream_t s) {
  if (cin == NET_C0)
    hipLaunchKernelGGL((k_reduce_conv_grad<NET_C0>), grid1(NET_C * (9 * NET_C0 + 1)), dim3(256), 0, s, nblocks,
                       part, gW, gB);
  else
    hipLaunchKernelGGL((k_reduce_conv_grad<NET_C>), grid1(NET_C * (9 * NET_C + 1)), dim3(256), 0, s, nblocks, part,
                       gW, gB);
}

void launch_reduce_head_grad(int nblocks, const float* part, float* grad, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_head_grad, grid1(9 * NET_C + 1), dim3(256), 0, s, nblocks, part, grad);
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

void launch_sum_loss(const double* part, int n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(256), 0, s, part, n, out);
}

void launch_features(const FeatIn& f, __bf16* X0, hipStream_t s) {
  const Img m{f.g.hi[1] - f.g.lo[1] + 1, f.g.hi[0] - f.g.lo[0] + 1};
  hipLaunchKernelGGL(k_features, grid1((int64_t)m.H * m.W), dim3(256), 0, s, f, m, X0);
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
