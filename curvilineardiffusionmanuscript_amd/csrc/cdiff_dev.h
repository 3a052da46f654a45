// cdiff_dev.h -- device-side types and helpers shared by the libcdiff kernels.
//
// Layout in HBM.  Every field is stored with a padded pitch: element (i, j, k)
// of a full (halo-inclusive) array lives at  off + i + pitch*j + plane*k,
// with off = ALIGN - nhalo so that the first domain cell of every row sits on
// a 128-byte boundary, and pitch a multiple of ALIGN.  Host arrays (Julia
// column-major, unpadded) are converted by pitched copies.
#pragma once
#include <type_traits>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cdiff {

constexpr int ALIGN = 16;  // elements: 128 B for FP64, 64 B for FP32
constexpr int WAVE = 64;
constexpr int OWN = WAVE - 2;  // columns owned by a 64-lane strip (lanes 1..62)

struct Geo {
  int ndim;
  int n[3];    // full dims (halo included)
  int lo[3];   // first domain index per axis
  int hi[3];   // last domain index per axis
  int64_t pitch;  // stride of j (elements)
  int64_t plane;  // stride of k (elements)
  int64_t off;    // offset of element (0,0,0)
  int64_t total;  // elements in an allocation
};

__host__ __device__ inline int64_t at(const Geo& g, int i, int j, int k) {
  return g.off + (int64_t)i + g.pitch * (int64_t)j + g.plane * (int64_t)k;
}

// Ghost-cell rule of one face of the domain (first halo layer), derived on the
// host from the BC tuple in application order (ilo, ihi, jlo, jhi, klo, khi):
//   COPY: u(src plane), NEG: -u(src plane), CONST: val, KEEP: stored value.
enum GhostMode { G_COPY = 0, G_CONST = 1, G_NEG = 2, G_KEEP = 3 };
struct Ghost {
  int mode[6];
  int src[6];     // source plane index along the face's axis
  double val[6];
};

// Convergence control block, one per solver, in device memory.
struct Ctrl {
  int done;        // 1 once the pseudo-iteration loop has ended
  int status;      // cdiff_status of the loop
  int final_iter;  // iteration at which it ended
  int nchecks;
  double rel, abs_, init;  // (inner correction loop: init = the smallest residual of its checks so far)
  double last_l2;
  int valid_flags;  // step-start validity (bit 0: alpha domain, bits 1..: halo sides; bit 16+: source)
  int grow;         // inner correction loop: consecutive checks whose residual grew
  int ref_iter;     // inner correction loop: iteration at which the residual last halved
  double ref;       //   and the residual it halved to
  int final_buf;    // inner correction loop: buffer written by the launch of its last check
};

// v[f] of a 6-entry table (the per-face BC rules of a Ghost) for a face index
// known only at run time, by selection: an indexed read would place the
// kernel argument in scratch memory
template <typename T>
__device__ __forceinline__ T pick6(const T (&v)[6], int f) {
  T r = v[0];
#pragma unroll
  for (int q = 1; q < 6; ++q) r = f == q ? v[q] : r;
  return r;
}

// Julia `x * isfinite(x)` (Bool strong zero).
template <typename T>
__device__ __forceinline__ T strong_zero(T x) {
  return isfinite(x) ? x : copysign(T(0), x);
}

template <typename T>
__device__ __forceinline__ T face_mean(int harmonic, T a, T b) {
  return harmonic ? (T(2) * a * b) / (a + b) : (a + b) / T(2);
}

// |x|^p with Julia literal powers for the integer exponents 0..3.
template <typename T>
__device__ __forceinline__ T lit_pow(T x, T p, int p_int) {
  if (p_int >= 0) {
    if (p_int == 0) return T(1);
    if (p_int == 1) return x;
    if (p_int == 2) return x * x;
    return x * x * x;
  }
  return pow(x, p);
}

struct KappaLaw {
  double kappa0, t_exp, rho_exp, cp;
  int t_int, rho_int;  // integer exponent 0..3, or -1 for pow()
  int t_half;          // t_exp = t_half + 1/2 (0..3): x^t_half * sqrt(x), or -1
};

// |x|^(n + 1/2) as a literal power times a square root: within 2 ulp of
// pow(x, n + 0.5) at a fraction of its cost.
template <typename T>
__device__ __forceinline__ T half_pow(T x, int n) {
  const T r = sqrt(x);
  if (n == 0) return r;
  if (n == 1) return x * r;
  if (n == 2) return (x * x) * r;
  return (x * x * x) * r;
}

// alpha = |kappa(rho, T) / (rho cp)| (src/pseudo_transient/conductivity.jl:7-10).
// UNIT: no density array and cp = 1, so rho cp = 1 and the division, which
// would return kappa exactly, is left out.
template <typename T, bool UNIT = false>
__device__ __forceinline__ T alpha_of(const KappaLaw& k, T rho, T u) {
  T kap;
  if (!isfinite(u)) {
    kap = T(0);
  } else {
    T tt = k.t_half >= 0 ? half_pow<T>(fabs(u), k.t_half) : lit_pow<T>(fabs(u), (T)k.t_exp, k.t_int);
    if (k.rho_exp == 0.0) {
      kap = (T)k.kappa0 * tt;
    } else {
      T rt = lit_pow<T>(fabs(rho), (T)k.rho_exp, k.rho_int);
      kap = (T)k.kappa0 * rt * tt;
    }
  }
  if (UNIT) return fabs(kap);
  return fabs(kap / (rho * (T)k.cp));
}

// Iteration parameters (src/pseudo_transient/iteration_parameters.jl:2-15).
struct IterConst {
  double L, Vpdtau, dt;
};

template <typename T>
__device__ __forceinline__ void iter_params(const IterConst& c, T rho, T a, T& dtau, T& theta) {
  const T pi = (T)3.141592653589793;
  const T L = (T)c.L, V = (T)c.Vpdtau, dt = (T)c.dt;
  T Re = pi + sqrt(pi * pi + (L * L * rho) / (a * dt));
  dtau = strong_zero<T>(V * L / (a * Re));
  theta = strong_zero<T>(L / (V * Re));
}

// Wave-wide neighbour exchange (lane l receives lane l-1 / l+1).
template <typename T>
__device__ __forceinline__ T from_left(T v) {
  return __shfl_up(v, 1, WAVE);
}
template <typename T>
__device__ __forceinline__ T from_right(T v) {
  return __shfl_down(v, 1, WAVE);
}

// NaN-propagating max for the Inf-norm flavour.
template <typename T>
__device__ __forceinline__ T nanmax(T a, T b) {
  if (a != a) return a;
  if (b != b) return b;
  return a > b ? a : b;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
  for (int o = WAVE / 2; o > 0; o >>= 1) v = nanmax(v, __shfl_xor(v, o, WAVE));
  return v;
}

// A marching loop unrolled over a ring period R: step(integral_constant<P>,
// t) for t = t, t + 1, ... with P = 0 .. R - 1, so that ring slots indexed by
// P are compile-time constants; false once t passes t1.
template <int P, int R, class F>
__device__ __forceinline__ bool march_phases(F& step, int& t, int t1) {
  if constexpr (P == R) {
    return true;
  } else {
    step(std::integral_constant<int, P>{}, t);
    if (++t > t1) return false;
    return march_phases<P + 1, R>(step, t, t1);
  }
}

}  // namespace cdiff
