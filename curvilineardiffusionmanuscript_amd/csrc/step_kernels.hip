// step_kernels.hip -- the passes at the two ends of a time step, fused.
//
// The logical elements of a full (halo-inclusive) array are those of the
// padded range off + e, e in [0, plane n2), whose column e mod pitch is below
// n0.  k_step_begin and k_step_end stride over that range with STEP_BLOCKS
// blocks, STEP_UNROLL independent elements in flight per thread, and leave a
// per-block NaN-propagating maximum (Julia `maximum`) in part[b]; k_dc_end
// takes one element per thread.  Fields hold fewer than 2^31 elements
// (cdiff_create), so 32-bit index arithmetic suffices.
#include "cdiff_dev.h"
#include "kernels.h"

namespace cdiff {

namespace {

constexpr int STEP_UNROLL = 8;

__device__ __forceinline__ double block_nanmax(double m, double* sm) {
  m = wave_max(m);
  if ((threadIdx.x & (WAVE - 1)) == 0) sm[threadIdx.x / WAVE] = m;
  __syncthreads();
  double r = -INFINITY;
  if (threadIdx.x == 0)
    for (int w = 0; w < STEP_THREADS / WAVE; ++w) r = nanmax(r, sm[w]);
  return r;
}

// validity bits of one value of alpha (or of the source) at (i, j, k):
// bit 0 domain, bits 1 + 2 ax / 2 + 2 ax the low / high halo side of axis ax,
// over expand(domain, 1) (validity_checks.jl:85-127)
__device__ __forceinline__ int validity_bits(const Geo& g, int i, int j, int k, double v, int positivity) {
  const int idx[3] = {i, j, k};
  bool inside = true;
  for (int ax = 0; ax < g.ndim; ++ax) inside = inside && idx[ax] >= g.lo[ax] - 1 && idx[ax] <= g.hi[ax] + 1;
  if (!inside) return 0;
  const bool bad = !isfinite(v) || (positivity && !(v >= 0.0));
  if (!bad) return 0;
  bool dom = true;
  for (int ax = 0; ax < g.ndim; ++ax) dom = dom && idx[ax] >= g.lo[ax] && idx[ax] <= g.hi[ax];
  int f = dom ? 1 : 0;
  for (int ax = 0; ax < g.ndim; ++ax) {
    if (idx[ax] == g.lo[ax] - 1) f |= 1 << (1 + 2 * ax);
    if (idx[ax] == g.hi[ax] + 1) f |= 1 << (2 + 2 * ax);
  }
  return f;
}

}  // namespace

// Start of a step (PseudoTransientSolver.jl:229-240): u0 = u1 = up = T; the
// validity flags of alpha (k_validate, shift 0) OR-ed into ctrl; and, when
// part is given, the per-block maximum of T, the first pass of next_dt
// (timestep_control.jl:45: maximum(u_prev)).
__global__ __launch_bounds__(STEP_THREADS) void k_step_begin(Geo g, const double* __restrict__ T,
                                                             double* __restrict__ u0, double* __restrict__ u1,
                                                             double* __restrict__ up, const double* __restrict__ alpha,
                                                             int positivity, Ctrl* ctrl, double* __restrict__ part) {
  __shared__ double sm[STEP_THREADS / WAVE];
  const uint32_t n = (uint32_t)(g.plane * g.n[2]), pitch = (uint32_t)g.pitch, n1 = (uint32_t)g.n[1];
  const uint32_t stride = gridDim.x * STEP_THREADS * STEP_UNROLL;
  double m = -INFINITY;
  int flags = 0;
  for (uint32_t base = blockIdx.x * STEP_THREADS * STEP_UNROLL; base < n; base += stride) {
    double v[STEP_UNROLL], a[STEP_UNROLL];
#pragma unroll
    for (int s = 0; s < STEP_UNROLL; ++s) {
      const uint32_t e = base + s * STEP_THREADS + threadIdx.x;
      v[s] = e < n ? T[g.off + e] : 0.0;
      a[s] = e < n ? alpha[g.off + e] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < STEP_UNROLL; ++s) {
      const uint32_t e = base + s * STEP_THREADS + threadIdx.x;
      const uint32_t row = e / pitch, i = e - row * pitch;
      if (e < n && i < (uint32_t)g.n[0]) {
        u0[g.off + e] = v[s];
        u1[g.off + e] = v[s];
        up[g.off + e] = v[s];
        m = nanmax(m, v[s]);
        flags |= validity_bits(g, (int)i, (int)(row % n1), (int)(row / n1), a[s], positivity);
      }
    }
  }
  for (int o = WAVE / 2; o > 0; o >>= 1) flags |= __shfl_xor(flags, o, WAVE);
  if ((threadIdx.x & (WAVE - 1)) == 0 && flags) atomicOr(&ctrl->valid_flags, flags);
  if (part) {
    const double b = block_nanmax(m, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = b;
  }
}

// End of a step (PseudoTransientSolver.jl:320-330, timestep_control.jl:33-65):
// T = u, with abs() under enforce_positivity and cutoff!; and, when part is
// given, the per-block maximum of the strong-zeroed relative change
// |T - up| / (T + umax tol), umax = maximum(up) from the start of the step.
__global__ __launch_bounds__(STEP_THREADS) void k_step_end(Geo g, const double* __restrict__ u, double* __restrict__ T,
                                                           int positivity, int cutoff, const double* __restrict__ up,
                                                           const double* __restrict__ umax, double tol,
                                                           double* __restrict__ part) {
  __shared__ double sm[STEP_THREADS / WAVE];
  const uint32_t n = (uint32_t)(g.plane * g.n[2]), pitch = (uint32_t)g.pitch;
  const uint32_t stride = gridDim.x * STEP_THREADS * STEP_UNROLL;
  const double u0 = part ? umax[0] * tol : 0.0;
  double m = -INFINITY;
  for (uint32_t base = blockIdx.x * STEP_THREADS * STEP_UNROLL; base < n; base += stride) {
    double v[STEP_UNROLL], p[STEP_UNROLL];
#pragma unroll
    for (int s = 0; s < STEP_UNROLL; ++s) {
      const uint32_t e = base + s * STEP_THREADS + threadIdx.x;
      v[s] = e < n ? u[g.off + e] : 0.0;
      p[s] = (part && e < n) ? up[g.off + e] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < STEP_UNROLL; ++s) {
      const uint32_t e = base + s * STEP_THREADS + threadIdx.x;
      if (e < n && e % pitch < (uint32_t)g.n[0]) {
        double a = v[s];
        if (positivity) a = fabs(a);
        if (cutoff) {
          const double h = 0.5 * (fabs(a) + a);
          a = isfinite(a) ? h : copysign(0.0, h);
        }
        T[g.off + e] = a;
        if (part) m = nanmax(m, strong_zero(fabs(a - p[s]) / (a + u0)));
      }
    }
  }
  if (part) {
    const double b = block_nanmax(m, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = b;
  }
}

// End of a defect-correction step, from its final FP64 iterate u: a copy of
// u, alpha(u) over the whole array (k_alpha_full) and the plain flux
// q' = -abar (u[+1] - u) of every axis on expand_lower(domain, axis), 0
// elsewhere (k_plain_flux), the APT's fixed point of the damped flux.  The
// neighbours' alpha is recomputed from their u (the same function, so the
// same bits as k_alpha_full's).  UNIT: no density array and cp = 1.
template <int NDIM, bool UNIT>
__global__ __launch_bounds__(STEP_THREADS) void k_dc_end(Geo g, KappaLaw kap, const double* __restrict__ rho,
                                                         int harmonic, const double* __restrict__ u,
                                                         double* __restrict__ ucopy, double* __restrict__ alpha,
                                                         double* __restrict__ q0, double* __restrict__ q1,
                                                         double* __restrict__ q2) {
  const uint32_t n = (uint32_t)(g.plane * g.n[2]), pitch = (uint32_t)g.pitch, n1 = (uint32_t)g.n[1];
  const uint32_t e = blockIdx.x * STEP_THREADS + threadIdx.x;
  if (e >= n) return;
  const uint32_t row = e / pitch, ii = e - row * pitch;
  if (ii >= (uint32_t)g.n[0]) return;
  const int idx[3] = {(int)ii, (int)(row % n1), (int)(row / n1)};
  const int64_t c = g.off + e;
  const int64_t st[3] = {1, g.pitch, g.plane};
  double* const q[3] = {q0, q1, q2};
  // faces computed: expand_lower(domain, ax) holds idx[b] in [lo_b - (b == ax), hi_b]
  bool in_b[3], lo_b[3];
#pragma unroll
  for (int b = 0; b < NDIM; ++b) {
    in_b[b] = idx[b] >= g.lo[b] && idx[b] <= g.hi[b];
    lo_b[b] = idx[b] == g.lo[b] - 1;
  }
  const double uc = u[c];
  double un[NDIM], rn[NDIM];
#pragma unroll
  for (int ax = 0; ax < NDIM; ++ax) {
    bool comp = true;
#pragma unroll
    for (int b = 0; b < NDIM; ++b) comp = comp && (in_b[b] || (b == ax && lo_b[b]));
    un[ax] = comp ? u[c + st[ax]] : 0.0;
    rn[ax] = (comp && rho) ? rho[c + st[ax]] : 1.0;
  }
  const double a = alpha_of<double, UNIT>(kap, rho ? rho[c] : 1.0, uc);
  ucopy[c] = uc;
  alpha[c] = a;
#pragma unroll
  for (int ax = 0; ax < NDIM; ++ax) {
    bool comp = true;
#pragma unroll
    for (int b = 0; b < NDIM; ++b) comp = comp && (in_b[b] || (b == ax && lo_b[b]));
    double f = 0.0;
    if (comp) {
      const double an = alpha_of<double, UNIT>(kap, rn[ax], un[ax]);
      f = -face_mean<double>(harmonic, a, an) * (un[ax] - uc);
    }
    q[ax][c] = f;
  }
}

void launch_dc_end(const Geo& g, const KappaLaw& kap, const double* rho, int harmonic, const double* u,
                   double* ucopy, double* alpha, double* const q[3], hipStream_t stream) {
  const int64_t n = g.plane * g.n[2];
  const dim3 grid((unsigned)((n + STEP_THREADS - 1) / STEP_THREADS));
  const bool unit = rho == nullptr && kap.cp == 1.0;
  if (g.ndim == 2) {
    if (unit)
      hipLaunchKernelGGL((k_dc_end<2, true>), grid, dim3(STEP_THREADS), 0, stream, g, kap, rho, harmonic, u, ucopy,
                         alpha, q[0], q[1], q[2]);
    else
      hipLaunchKernelGGL((k_dc_end<2, false>), grid, dim3(STEP_THREADS), 0, stream, g, kap, rho, harmonic, u, ucopy,
                         alpha, q[0], q[1], q[2]);
  } else {
    if (unit)
      hipLaunchKernelGGL((k_dc_end<3, true>), grid, dim3(STEP_THREADS), 0, stream, g, kap, rho, harmonic, u, ucopy,
                         alpha, q[0], q[1], q[2]);
    else
      hipLaunchKernelGGL((k_dc_end<3, false>), grid, dim3(STEP_THREADS), 0, stream, g, kap, rho, harmonic, u, ucopy,
                         alpha, q[0], q[1], q[2]);
  }
}

}  // namespace cdiff
