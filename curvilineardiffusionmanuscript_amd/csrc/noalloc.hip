// noalloc.hip -- next_dt on device arrays with no allocation: the reduction
// partials live in a file-scope __device__ array, one copy per GPU, whose
// address each device resolves once with hipGetSymbolAddress.
//
// next_dt(u_new, u_prev, dt; tol, maximum_u_change, timestep_growth_factor)
// (src/timestep_control.jl:33-54) over every element of the arrays (halo and
// corners included, as the reference's full-array max):
//   umax = max(u_prev); u0 = umax * tol
//   delta = |u_new - u_prev| / (u_new + u0), 0 where not finite
//   dt' = dt * sqrt(|maximum_u_change / max(delta)|); return min(dt', growth dt)
// The same arithmetic as cdiff_next_dt's kernels; a maximum does not depend on
// the order of its operands, so the two give the same bits.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "../../include/cdiff.h"
#include "cdiff_dev.h"

using namespace cdiff;

namespace {

constexpr int ND_THREADS = 256;
constexpr int ND_PER_THREAD = 16;
constexpr int64_t ND_CELLS_PER_BLOCK = (int64_t)ND_THREADS * ND_PER_THREAD;  // 4096
constexpr int ND_MAX_PARTIALS = 32768;
// the scratch: one partial maximum per block, then [umax, max delta]
constexpr int ND_SCRATCH = ND_MAX_PARTIALS + 2;
constexpr int ND_MAX_DEVICES = 64;

}  // namespace

// per-GPU scratch (256 KiB + 16 B), held until the process ends
static __device__ __align__(256) double scratch[ND_SCRATCH];

namespace {

std::mutex g_mu;
double* g_scratch[ND_MAX_DEVICES] = {nullptr};

// pass 0: partial max of u_prev; pass 1: partial max of the relative change,
// with u0 = tol * (the pass-0 maximum in red[0])
__global__ __launch_bounds__(ND_THREADS) void k_nd_partial(const double* __restrict__ unew,
                                                           const double* __restrict__ uold, int64_t n, double tol,
                                                           int pass, double* __restrict__ part,
                                                           const double* __restrict__ red) {
  const double u0 = pass ? red[0] * tol : 0.0;
  const int64_t base = (int64_t)blockIdx.x * ND_CELLS_PER_BLOCK + threadIdx.x;
  double m = -INFINITY;
#pragma unroll 4
  for (int r = 0; r < ND_PER_THREAD; ++r) {
    const int64_t e = base + (int64_t)r * ND_THREADS;
    if (e < n) {
      double v;
      if (pass == 0) {
        v = uold[e];
      } else {
        const double a = unew[e];
        v = strong_zero(fabs(a - uold[e]) / (a + u0));
      }
      m = nanmax(m, v);
    }
  }
  m = wave_max(m);
  __shared__ double sm[ND_THREADS / WAVE];
  if ((threadIdx.x & (WAVE - 1)) == 0) sm[threadIdx.x / WAVE] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = sm[0];
    for (int w = 1; w < ND_THREADS / WAVE; ++w) b = nanmax(b, sm[w]);
    part[blockIdx.x] = b;
  }
}

// max of `np` partials into *out; with `final`, also the next time step
__global__ __launch_bounds__(1024) void k_nd_reduce(const double* __restrict__ part, int np, double* __restrict__ out,
                                                    int final, double dt, double muc, double growth,
                                                    double* __restrict__ dt_out) {
  double m = -INFINITY;
  for (int p = threadIdx.x; p < np; p += 1024) m = nanmax(m, part[p]);
  m = wave_max(m);
  __shared__ double sm[1024 / WAVE];
  if ((threadIdx.x & (WAVE - 1)) == 0) sm[threadIdx.x / WAVE] = m;
  __syncthreads();
  if (threadIdx.x != 0) return;
  double b = sm[0];
  for (int w = 1; w < 1024 / WAVE; ++w) b = nanmax(b, sm[w]);
  out[0] = b;
  if (!final) return;
  const double dtn = dt * sqrt(fabs(muc / b));
  const double gdt = growth * dt;
  double r;
  if (dtn != dtn)
    r = dtn;
  else if (gdt != gdt)
    r = gdt;
  else
    r = dtn < gdt ? dtn : gdt;
  dt_out[0] = r;
}

void put_err(char* buf, size_t len, const char* msg) {
  if (buf && len > 0) std::snprintf(buf, len, "%s", msg);
}

}  // namespace

extern "C" int cdiff_next_dt_device(const double* u_new, const double* u_prev, int64_t n, double dt, double tol,
                                    double maximum_u_change, double timestep_growth_factor, double* out,
                                    void* stream, char* errbuf, size_t errlen) {
  if (!u_new || !u_prev || !out || n < 1) {
    put_err(errbuf, errlen, "cdiff_next_dt_device: null array or n < 1");
    return CDIFF_BAD_ARG;
  }
  const int64_t nblk = (n + ND_CELLS_PER_BLOCK - 1) / ND_CELLS_PER_BLOCK;
  if (nblk > ND_MAX_PARTIALS) {
    char m[192];
    std::snprintf(m, sizeof m,
                  "cdiff_next_dt_device: %lld elements exceed the scratch of %d partials (at most %lld elements)",
                  (long long)n, ND_MAX_PARTIALS, (long long)ND_MAX_PARTIALS * ND_CELLS_PER_BLOCK);
    put_err(errbuf, errlen, m);
    return CDIFF_BAD_ARG;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= ND_MAX_DEVICES) {
    put_err(errbuf, errlen, "cdiff_next_dt_device: no current HIP device");
    return CDIFF_HIP_ERROR;
  }
  double* s = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
      // this device's copy of the scratch array: its address, resolved once
      void* p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) {
        put_err(errbuf, errlen, "cdiff_next_dt_device: hipGetSymbolAddress(scratch) failed");
        return CDIFF_HIP_ERROR;
      }
      g_scratch[dev] = static_cast<double*>(p);
    }
    s = g_scratch[dev];
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  double* part = s;
  double* red = s + ND_MAX_PARTIALS;
  const int np = (int)nblk;
  hipLaunchKernelGGL(k_nd_partial, dim3((unsigned)np), dim3(ND_THREADS), 0, st, u_new, u_prev, n, tol, 0, part,
                     (const double*)red);
  hipLaunchKernelGGL(k_nd_reduce, dim3(1), dim3(1024), 0, st, (const double*)part, np, red, 0, dt, maximum_u_change,
                     timestep_growth_factor, out);
  hipLaunchKernelGGL(k_nd_partial, dim3((unsigned)np), dim3(ND_THREADS), 0, st, u_new, u_prev, n, tol, 1, part,
                     (const double*)red);
  hipLaunchKernelGGL(k_nd_reduce, dim3(1), dim3(1024), 0, st, (const double*)part, np, red + 1, 1, dt,
                     maximum_u_change, timestep_growth_factor, out);
  e = hipGetLastError();
  if (e != hipSuccess) {
    put_err(errbuf, errlen, hipGetErrorString(e));
    return CDIFF_HIP_ERROR;
  }
  return CDIFF_OK;
}
