// solver.cpp -- the time step of libcdiff (PseudoTransientScheme.step!,
// src/pseudo_transient/PseudoTransientSolver.jl:180-351) on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "kernels.h"
#include "solver.h"
#include "net.h"
#include "sweep_args.h"

using namespace cdiff;

#define HIP_TRY(expr)                                  \
  do {                                                 \
    hipError_t _e = (expr);                            \
    if (_e != hipSuccess) return hip_fail(_e, #expr);  \
  } while (0)

#define LAUNCH_CHECK(name)                                          \
  do {                                                              \
    hipError_t _e = hipGetLastError();                              \
    if (_e != hipSuccess) return hip_fail(_e, "launch of " name);   \
  } while (0)

namespace {

dim3 grid_full(const Geo& g) {
  const int64_t n = (int64_t)g.n[0] * g.n[1] * g.n[2];
  return dim3((unsigned)((n + 255) / 256));
}

// The error-check schedule of PseudoTransientSolver.jl:206-208 and :280:
// m = round(0.9 N) (ties to even), k = max(1, (N - m) / 2).
void schedule(int last, int& m, int& k) {
  m = (int)std::nearbyint((double)last * 0.9);
  k = std::max(1, (last - m) / 2);
}

// A Float64 as Julia's show prints it (the text of the reference's error
// messages, e.g. "$((_min, _max))", PseudoTransientSolver.jl:302-314,
// validity_checks.jl:14-17): the shortest digits that read back as the same
// double; fixed notation when the decimal point falls at -3..6 places
// (0.0001, 123456.0), scientific otherwise (1.0e-5, 1.0e6); "Inf", "NaN",
// and ".0" on integral values.
std::string julia_float(double x) {
  if (std::isnan(x)) return "NaN";
  if (std::isinf(x)) return x < 0 ? "-Inf" : "Inf";
  const std::string sign = std::signbit(x) ? "-" : "";
  const double ax = std::fabs(x);
  if (ax == 0.0) return sign + "0.0";
  char e[40];
  for (int p = 1; p <= 17; ++p) {
    std::snprintf(e, sizeof e, "%.*e", p - 1, ax);
    if (std::strtod(e, nullptr) == ax) break;
  }
  // ax = d1.d2..dn x 10^x10 = 0.d1..dn x 10^pt
  std::string dig;
  const char* c = e;
  for (; *c && *c != 'e'; ++c)
    if (*c >= '0' && *c <= '9') dig += *c;
  while (dig.size() > 1 && dig.back() == '0') dig.pop_back();
  const int n = (int)dig.size();
  const int x10 = std::atoi(c + 1), pt = x10 + 1;
  if (pt > -4 && pt <= 6) {
    if (pt <= 0) return sign + "0." + std::string(-pt, '0') + dig;
    if (pt < n) return sign + dig.substr(0, pt) + "." + dig.substr(pt);
    return sign + dig + std::string(pt - n, '0') + ".0";
  }
  return sign + dig.substr(0, 1) + "." + (n > 1 ? dig.substr(1) : std::string("0")) + "e" + std::to_string(x10);
}

inline bool is_check(int it, int m, int k) { return (it >= m && it % k == 0) || it == 1; }

}  // namespace

int cdiff_solver::ensure_parts() {
  if (nparts <= part_cap) return CDIFF_OK;
  double* ps = nullptr;
  double* pm = nullptr;
  HIP_TRY(hipMalloc((void**)&ps, sizeof(double) * nparts));
  HIP_TRY(hipMalloc((void**)&pm, sizeof(double) * nparts));
  for (double* old : {part_sum, part_max})
    if (old) {
      HIP_TRY(hipStreamSynchronize(stream));
      owned.erase(std::remove(owned.begin(), owned.end(), (void*)old), owned.end());
      HIP_TRY(hipFree(old));
    }
  part_sum = ps;
  part_max = pm;
  owned.push_back(ps);
  owned.push_back(pm);
  part_cap = nparts;
  return CDIFF_OK;
}

void cdiff_solver::configure_sweep() {
  tile2d = tune.sweep2d == 0;
  const int ni = g.hi[0] - g.lo[0] + 1;
  const int nj = g.hi[1] - g.lo[1] + 1;
  if (ndim == 2 && tile2d) {
    // LDS-tiled kernel: 64 x 4 cells per block
    strips = (ni + 63) / 64;
    const int ty = tune.tile_rows;
    ch = ty;
    chunks = (nj + ty - 1) / ty;
    ntiles = strips * chunks;
    nblocks = (ntiles + 7) / 8 * 8;
  } else if (ndim == 2) {
    strips = (ni + OWN - 1) / OWN;
    int c = tune.march_rows64;
    if (c <= 0) {
      // aim for ~4096 waves (16 per CU), rows per chunk between 8 and 64
      const int64_t cells = (int64_t)strips * nj;
      c = (int)std::min<int64_t>(64, std::max<int64_t>(8, (cells + 4095) / 4096));
    }
    ch = std::min(c, nj);
    chunks = (nj + ch - 1) / ch;
    ntiles = strips * chunks;
    nblocks = (ntiles + 7) / 8 * 8;
  } else {
    const int nk = g.hi[2] - g.lo[2] + 1;
    col3d = tune.sweep3d == 0;
    int c = tune.planes3d;
    if (col3d) {
      strips3[0] = (ni + 63) / 64;
      strips3[1] = (nj + 7) / 8;
      if (c <= 0) {
        // one 512-thread block per CU (its LDS and registers allow one at a
        // time): long columns, fewer ring planes; 4..256 planes per chunk
        // (256^3: 128 planes, 822 us per sweep against 871 us at 32)
        const int64_t cols = (int64_t)strips3[0] * strips3[1];
        c = (int)std::min<int64_t>(256, std::max<int64_t>(4, (cols * nk + ncu - 1) / ncu));
      }
    } else {
      strips3[0] = (ni + OWN - 1) / OWN;
      strips3[1] = (nj + (SWEEP3D_BJ - 2) - 1) / (SWEEP3D_BJ - 2);
      if (c <= 0) {
        // aim for ~2048 blocks of SWEEP3D_BJ waves, 4..64 planes per chunk
        const int64_t cols = (int64_t)strips3[0] * strips3[1];
        c = (int)std::min<int64_t>(64, std::max<int64_t>(4, (cols * nk + 2047) / 2048));
      }
    }
    c = std::min(c, nk);
    ch = c;
    chunks = (nk + ch - 1) / ch;
    ntiles = strips3[0] * strips3[1] * chunks;
    nblocks = (ntiles + 7) / 8 * 8;
  }
  nparts = nblocks;
}

bool cdiff_solver::pairs_ok() const {
  if (ndim != 2 || !tile2d || ch != 8 || !tune.pairs) return false;
  for (int f = 0; f < 4; ++f)
    if (bcs[f].kind == CDIFF_BC_PERIODIC) return false;
  return true;
}

Sweep2DArgs<double> cdiff_solver::sweep2d_args(const cdiff_opts& o, double dt, double Vpdtau, int src) const {
  const int dst = 1 - src;
  Sweep2DArgs<double> A{};
  A.g = g;
  A.gh = gh;
  A.kap = kap;
  A.ic.L = L;
  A.ic.Vpdtau = Vpdtau;
  A.ic.dt = dt;
  A.harmonic = mean == CDIFF_MEAN_HARMONIC;
  A.ntiles = ntiles;
  A.chunks = chunks;
  A.ch = ch;
  A.u_src = u[src];
  A.u_dst = u[dst];
  A.qx_src = q[0][src];
  A.qx_dst = q[0][dst];
  A.qy_src = q[1][src];
  A.qy_dst = q[1][dst];
  A.u_prev = u_prev;
  A.g11 = coef[0];
  A.g22 = coef[1];
  A.c12 = coef[2];
  A.ac = coef[3];
  A.bc = coef[4];
  A.S = S;
  A.rho = rho;
  A.tzA = tzA;
  A.tzB = tzB;
  if (!o.subcycle_conductivity) {
    A.frozen_alpha = frozen[0];
    A.frozen_theta = frozen[1];
    A.frozen_dtau = frozen[2];
  }
  A.part_sum = part_sum;
  A.part_max = part_max;
  A.ctrl = ctrl;
  return A;
}

int cdiff_solver::launch_sweep64(const cdiff_opts& o, double dt, double Vpdtau, int src, bool check, bool pair) {
  if (ndim == 2) {
    Sweep2DArgs<double> A = sweep2d_args(o, dt, Vpdtau, src);
    if (graph_capture) A.ic_dev = ic_dev;
    if (pair) {
      A.u_pre = check ? u_pre : nullptr;
      launch_sweep2d_pair<double>(A, check, nblocks, stream);
    } else if (tile2d) {
      launch_sweep2d_tile<double>(A, check, nblocks, stream);
    } else {
      launch_sweep2d<double>(A, check, nblocks, stream);
    }
    LAUNCH_CHECK("sweep2d");
    return CDIFF_OK;
  }
  const int dst = 1 - src;
  Sweep3DArgs<double> A{};
  A.g = g;
  A.gh = gh;
  A.kap = kap;
  A.ic.L = L;
  A.ic.Vpdtau = Vpdtau;
  A.ic.dt = dt;
  A.harmonic = mean == CDIFF_MEAN_HARMONIC;
  A.ntiles = ntiles;
  A.chunks = chunks;
  A.ch = ch;
  A.stripsI = strips3[0];
  A.u_src = u[src];
  A.u_dst = u[dst];
  A.qx_src = q[0][src];
  A.qx_dst = q[0][dst];
  A.qy_src = q[1][src];
  A.qy_dst = q[1][dst];
  A.qz_src = q[2][src];
  A.qz_dst = q[2][dst];
  A.u_prev = u_prev;
  for (int t = 0; t < 9; ++t) A.coef[t] = coef[t];
  A.S = S;
  A.rho = rho;
  A.tzA = tzA;
  A.tzB = tzB;
  if (!o.subcycle_conductivity) {
    A.frozen_alpha = frozen[0];
    A.frozen_theta = frozen[1];
    A.frozen_dtau = frozen[2];
  }
  A.part_sum = part_sum;
  A.part_max = part_max;
  A.ctrl = ctrl;
  A.coef_stride = g.total;
  if (col3d)
    launch_sweep3d_col<double>(A, check, nblocks, stream);
  else
    launch_sweep3d<double>(A, check, nblocks, stream);
  LAUNCH_CHECK("sweep3d");
  return CDIFF_OK;
}

int cdiff_solver::enqueue_next_dt(const double* unew, const double* uold, double dt, double tol, double muc,
                                  double growth) {
  double* part = red + 8;
  const int nb = 1024;
  hipLaunchKernelGGL(k_next_dt_partial, dim3(nb), dim3(256), 0, stream, g, unew, uold, (const double*)red, tol, 0,
                     part);
  hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(256), 0, stream, (const double*)part, nb, red + 0);
  hipLaunchKernelGGL(k_next_dt_partial, dim3(nb), dim3(256), 0, stream, g, unew, uold, (const double*)red, tol, 1,
                     part);
  hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(256), 0, stream, (const double*)part, nb, red + 1);
  hipLaunchKernelGGL(k_next_dt_final, dim3(1), dim3(1), 0, stream, (const double*)(red + 1), dt, muc, growth, red + 2);
  LAUNCH_CHECK("next_dt kernels");
  return CDIFF_OK;
}

bool cdiff_solver::graph_ok() const { return tune.graphs && ndim == 2 && tile2d && ic_dev; }

void cdiff_solver::graph_reset() {
  for (auto& b : gblock) {
    if (b.exec) (void)hipGraphExecDestroy(b.exec);
    b.exec = nullptr;
    b.key.clear();
  }
}

// GRAPH_ITERS pseudo-iterations without a check from buffer `src` (an even
// number of launches, so the block ends on the parity it started from): the
// cached graph when it was captured from the same launches, else a new
// capture.  The kernels read this step's dt and Vpdtau from ic_dev.
int cdiff_solver::graph_block(const cdiff_opts& o, double dt, double Vpdtau, int src, bool pair) {
  GraphBlock& gb = gblock[src];
  // everything the captured launches bake in besides the iteration constants:
  // buffers, optional arrays, the conductivity law and the launch shape (the
  // geometry and ghost rules are fixed for the handle)
  const Sweep2DArgs<double> A = sweep2d_args(o, 0.0, 0.0, src);
  auto bits = [](double v) {
    uint64_t b;
    std::memcpy(&b, &v, sizeof b);
    return b;
  };
  auto ptr = [](const void* p) { return (uint64_t)(uintptr_t)p; };
  const std::vector<uint64_t> key = {
      ptr(A.u_src), ptr(A.u_dst), ptr(A.qx_src), ptr(A.qx_dst), ptr(A.qy_src), ptr(A.qy_dst), ptr(A.u_prev),
      ptr(A.g11), ptr(A.g22), ptr(A.c12), ptr(A.ac), ptr(A.bc), ptr(A.S), ptr(A.rho), ptr(A.tzA), ptr(A.tzB),
      ptr(A.frozen_alpha), ptr(A.frozen_theta), ptr(A.frozen_dtau), ptr(A.part_sum), ptr(A.part_max),
      ptr(A.ctrl), ptr(ic_dev), bits(kap.kappa0), bits(kap.t_exp), bits(kap.rho_exp), bits(kap.cp),
      (uint64_t)kap.t_int, (uint64_t)kap.rho_int, (uint64_t)kap.t_half, (uint64_t)A.harmonic,
      (uint64_t)A.ntiles, (uint64_t)A.chunks, (uint64_t)A.ch, (uint64_t)nblocks, (uint64_t)pair,
      (uint64_t)tile2d};
  if (!gb.exec || gb.key != key) {
    if (gb.exec) (void)hipGraphExecDestroy(gb.exec);
    gb.exec = nullptr;
    gb.key.clear();
    const int nlaunch = pair ? GRAPH_ITERS / 2 : GRAPH_ITERS;
    HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    graph_capture = true;
    int rc = CDIFF_OK, b = src;
    for (int l = 0; l < nlaunch && rc == CDIFF_OK; ++l, b ^= 1) rc = launch_sweep64(o, dt, Vpdtau, b, false, pair);
    graph_capture = false;
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(stream, &graph);
    if (rc != CDIFF_OK || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc != CDIFF_OK ? rc : hip_fail(e, "hipStreamEndCapture");
    }
    const hipError_t ei = hipGraphInstantiate(&gb.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      gb.exec = nullptr;
      return hip_fail(ei, "hipGraphInstantiate");
    }
    gb.key = key;
  }
  HIP_TRY(hipGraphLaunch(gb.exec, stream));
  return CDIFF_OK;
}

// Launch the FP64 sweeps of one time step in segments (see cdiff.cpp).
int cdiff_solver::run_loop64(const cdiff_opts& o, double dt, double Vpdtau, int /*first_iter_base*/) {
  int m, k;
  schedule(last_iter, m, k);
  const int lmin = tune.seg_min;
  const int lmax = tune.seg_max;
  const int64_t limit = (int64_t)o.max_iter + 1;

  const bool pairs = pairs_ok();
  if (pairs && !u_pre) {
    HIP_TRY(hipMalloc((void**)&u_pre, sizeof(double) * g.total));
    owned.push_back(u_pre);
  }
  int64_t it = 0;
  int seg = 0;
  int buf = cur;  // source buffer of the next launch
  bool pending = false;
  launches.clear();
  const bool graphs = graph_ok();
  if (graphs) {
    *ic_host = IterConst{L, Vpdtau, dt};
    HIP_TRY(hipMemcpyAsync(ic_dev, ic_host, sizeof(IterConst), hipMemcpyHostToDevice, stream));
  }
  while (true) {
    const int64_t s = it + 1;
    const int64_t cap = std::min<int64_t>(s + lmax - 1, limit);
    int64_t e = std::min<int64_t>(s + lmin - 1, cap);
    while (e < cap && !is_check((int)e, m, k)) ++e;
    for (int64_t t = s; t <= e;) {
      if (graphs && t + GRAPH_ITERS - 1 <= e) {
        bool clear = true;
        for (int64_t x = t; x < t + GRAPH_ITERS && clear; ++x) clear = !is_check((int)x, m, k);
        if (clear) {
          int rc = graph_block(o, dt, Vpdtau, buf, pairs);
          if (rc) return rc;
          for (int l = 0, b = buf; l < (pairs ? GRAPH_ITERS / 2 : GRAPH_ITERS); ++l, b ^= 1)
            launches.push_back({(int)(t - 1 + (pairs ? 2 * (l + 1) : l + 1)), 1 - b, pairs});
          t += GRAPH_ITERS;
          continue;
        }
      }
      const bool chk = is_check((int)t, m, k);
      const bool pair = pairs && !chk && t + 1 <= e;
      const int64_t tl = pair ? t + 1 : t;  // last iteration of this launch
      const bool chk_last = pair ? is_check((int)tl, m, k) : chk;
      int rc = launch_sweep64(o, dt, Vpdtau, buf, chk_last, pair);
      if (rc) return rc;
      launches.push_back({(int)tl, 1 - buf, pair});
      buf ^= 1;
      if (chk_last || tl == limit) {
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, stream, ctrl, (const double*)part_sum,
                           (const double*)part_max, chk_last ? nparts : 0, (int)tl, o.max_iter, o.rtol, o.atol,
                           N_full, o.norm_kind, chk_last ? 1 : 0);
        LAUNCH_CHECK("k_finalize");
      } else if (pair && tl > limit) {
        // cannot happen: pairs end at or before the segment end, which is capped at the limit
      }
      t = tl + 1;
    }
    it = e;
    HIP_TRY(hipMemcpyAsync(&ctrl_h[seg & 1], ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipEventRecord(ev[seg & 1], stream));
    if (pending) {
      HIP_TRY(hipEventSynchronize(ev[(seg - 1) & 1]));
      if (ctrl_h[(seg - 1) & 1].done) break;
    }
    if (it >= limit) break;
    pending = true;
    ++seg;
  }
  HIP_TRY(hipMemcpyAsync(&ctrl_h[2], ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  return CDIFF_OK;
}

int cdiff_solver::validation_message(int status) {
  const bool is_alpha = status == CDIFF_INVALID_ALPHA;
  const double* dev = is_alpha ? alpha : S;
  const char* name = is_alpha ? "diffusivity" : "source_term";
  const int positivity = is_alpha ? enforce_pos_last : 0;
  const int64_t n = (int64_t)g.n[0] * g.n[1] * g.n[2];
  std::vector<double> a(n, 0.0);
  if (dev) {
    hipError_t e = hipMemcpy2D(a.data(), sizeof(double) * g.n[0], dev + g.off, sizeof(double) * g.pitch,
                               sizeof(double) * g.n[0], (size_t)g.n[1] * g.n[2], hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy2D(validate)");
  }
  auto idx = [&](int i, int j, int k) { return (int64_t)i + (int64_t)g.n[0] * (j + (int64_t)g.n[1] * k); };
  auto bad = [&](double v) { return !std::isfinite(v) || (positivity && !(v >= 0)); };
  bool dom_bad = false, any_nan = false;
  double mn = INFINITY, mx = -INFINITY;
  for (int k = g.lo[2]; k <= g.hi[2]; ++k)
    for (int j = g.lo[1]; j <= g.hi[1]; ++j)
      for (int i = g.lo[0]; i <= g.hi[0]; ++i) {
        const double v = a[idx(i, j, k)];
        dom_bad |= bad(v);
        any_nan |= std::isnan(v);
        mn = std::min(mn, v);
        mx = std::max(mx, v);
      }
  char buf[256];
  if (dom_bad) {
    if (any_nan) mn = mx = NAN;
    err = std::string("Invalid ") + name + " in the domain extrema -> (" + julia_float(mn) + ", " + julia_float(mx) +
          ")";
    return status;
  }
  int flo[3], fhi[3];
  for (int ax = 0; ax < 3; ++ax) {
    flo[ax] = ax < ndim ? g.lo[ax] - 1 : 0;
    fhi[ax] = ax < ndim ? g.hi[ax] + 1 : 0;
  }
  for (int ax = 0; ax < ndim; ++ax)
    for (int side = 0; side < 2; ++side) {
      const int plane = side ? fhi[ax] : flo[ax];
      bool b = false;
      for (int k = flo[2]; k <= fhi[2]; ++k)
        for (int j = flo[1]; j <= fhi[1]; ++j)
          for (int i = flo[0]; i <= fhi[0]; ++i) {
            const int c[3] = {i, j, k};
            if (c[ax] == plane && bad(a[idx(i, j, k)])) b = true;
          }
      if (b) {
        std::snprintf(buf, sizeof buf, "Invalid %s in the %s halo region for axis: %d", name, side ? "hi" : "lo",
                      ax + 1);
        err = buf;
        return status;
      }
    }
  err = std::string("Invalid ") + name;
  return status;
}

int cdiff_solver::step(double dt, const cdiff_opts& o, cdiff_stats* stats, double* next_dt_out) {
  const auto t0 = std::chrono::steady_clock::now();
  t_written = false;
  if (!metrics_set) return fail(CDIFF_NOT_READY, "cdiff_set_metrics has not been called");
  if (!(dt > 0)) return fail(CDIFF_BAD_ARG, "AssertionError: dt > 0");
  double smin = spacing[0];
  for (int a = 0; a < ndim; ++a) {
    if (!(spacing[a] > 0)) return fail(CDIFF_BAD_ARG, "AssertionError: all(solver.spacing .> 0)");
    smin = std::min(smin, spacing[a]);
  }
  const double Vpdtau = o.cfl * smin;
  if (!(Vpdtau > 0)) return fail(CDIFF_BAD_ARG, "AssertionError: Vpdτ > 0");
  if (o.max_iter < 0) return fail(CDIFF_BAD_ARG, "max_iter must be >= 0");
  if (o.norm_kind != 0 && o.norm_kind != 1) return fail(CDIFF_BAD_ARG, "norm_kind must be 0 or 1");
  enforce_pos_last = o.enforce_positivity;
  last_dt = dt;
  last_Vpdtau = Vpdtau;
  last_subcycle = o.subcycle_conductivity;
  cdiff_stats st{};

  // learned warm start: features of (T^n, T^{n-1}) and the prediction, before
  // the prologue overwrites u_prev
  warm_pending = false;
  warm_r0_ready = false;
  // (from the second step on: the network reads the previous increment
  // T^n - T^{n-1} and was trained on steps that have one)
  const bool warm_here = net && have_prev && (o.precision == CDIFF_PREC_FP32_DC || o.precision == CDIFF_PREC_FP64) &&
                         ndim == 2 && lin_ok();
  // after consultations that did not pay (warm_account), the network sits
  // out 15, 31, ... (at most 255) steps before it is consulted again; and it
  // is consulted only where the saving expected from it (the last step's work
  // per decade of residual reduction times the decades a prediction takes
  // off) exceeds what a consultation costs
  const bool always = tune.warm_policy == 1;
  const bool warm_skipped = warm_here && !always && warm_skip > 0;
  if (warm_skipped) --warm_skip;
  const bool consult = warm_here && !warm_skipped && (always || warm_pays());
  st.warm_consulted = consult ? 1 : 0;
  if (consult) {
    if (!learn_ts) {
      HIP_TRY(hipMalloc((void**)&learn_ts, 2 * sizeof(double)));
      owned.push_back(learn_ts);
    }
    int rc = net->workspace(Img{g.hi[1] - g.lo[1] + 1, g.hi[0] - g.lo[0] + 1});
    if (rc) return fail(rc, net->err);
    if ((rc = learn_features(dt, net->XS, learn_ts))) return rc;
    if ((rc = net->forward(net->XS, stream))) return fail(rc, net->err);
    warm_pending = true;
    // the feature pass evaluated the residual of T^n (u = u_prev = T^n, BC
    // ghosts by the pass's own rule) with the 2D outer kernel; nothing before
    // the step's first residual touches its partial sums unless the deeper
    // Dirichlet halo layers are filled first
    warm_r0_ready = lin && !dirichlet_outer;
  }

  // prologue: u = u_prev = T (PseudoTransientSolver.jl:229-230), validity
  // (:238-250) and maximum(u_prev) for next_dt, in one pass over T and alpha
  const int nrows = g.n[1] * g.n[2];
  const int step_blocks = std::min(nrows, STEP_BLOCKS);
  auto begin_step = [&]() {
    hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(1), 0, stream, ctrl, o.rtol, o.atol);
    hipLaunchKernelGGL(k_step_begin, dim3(step_blocks), dim3(STEP_THREADS), 0, stream, g, (const double*)Tdev, u[0],
                       u[1], u_prev, (const double*)alpha, o.enforce_positivity, ctrl,
                       o.calculate_next_dt ? red + 8 : nullptr);
    if (o.calculate_next_dt)
      hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(256), 0, stream, (const double*)(red + 8), step_blocks, red + 0);
    if (S) hipLaunchKernelGGL(k_validate, grid_full(g), dim3(256), 0, stream, g, (const double*)S, 0, 16, ctrl);
    hipLaunchKernelGGL(k_ctrl_validated, dim3(1), dim3(1), 0, stream, ctrl);
  };
  // the control block and validity again, the buffers kept (warm start)
  auto reset_and_validate = [&]() {
    hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(1), 0, stream, ctrl, o.rtol, o.atol);
    hipLaunchKernelGGL(k_validate, grid_full(g), dim3(256), 0, stream, g, (const double*)alpha, o.enforce_positivity,
                       0, ctrl);
    if (S) hipLaunchKernelGGL(k_validate, grid_full(g), dim3(256), 0, stream, g, (const double*)S, 0, 16, ctrl);
    hipLaunchKernelGGL(k_ctrl_validated, dim3(1), dim3(1), 0, stream, ctrl);
  };
  begin_step();
  LAUNCH_CHECK("prologue");
  if (dirichlet_outer)
    for (int b = 0; b < 2; ++b) {
      hipLaunchKernelGGL(k_halo_fill, grid_full(g), dim3(256), 0, stream, g, gh, u[b], 1);
      LAUNCH_CHECK("k_halo_fill");
    }

  bool fallback = false;
  if (o.precision == CDIFF_PREC_FP32_DC) {
    int rc = step_fp32_dc(dt, o, Vpdtau, &st);
    if (rc) return rc;
    if (dc_failed) {
      // the defect correction gave up: redo the step with the FP64 APT from
      // T^n (Tdev, u_prev and the FP64 fluxes are untouched by the correction
      // loop), so the step ends like an FP64 step would
      if (tune.verbose)
        std::fprintf(stderr, "[dc] defect correction gave up after %d rounds; FP64 APT from T^n\n", st.outer);
      fallback = true;
      begin_step();
      if (dirichlet_outer)
        for (int b = 0; b < 2; ++b) hipLaunchKernelGGL(k_halo_fill, grid_full(g), dim3(256), 0, stream, g, gh, u[b], 1);
      LAUNCH_CHECK("fallback prologue");
    }
  }
  if (o.precision == CDIFF_PREC_FP64 || fallback) {
    if (warm_pending) {
      // learned warm start of the FP64 APT (valid inputs only): the start with
      // the smaller FP64 residual, prediction or T^n, with its BC ghosts, in
      // both buffers; then the convergence control starts afresh
      HIP_TRY(hipMemcpyAsync(&ctrl_h[2], ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
      HIP_TRY(hipStreamSynchronize(stream));
      if (!ctrl_h[2].done) {
        int rc = ensure_fp32();
        if (rc) return rc;
        dc_Vpdtau = Vpdtau;
        if ((rc = warm_r0_ready ? dc_finalize_outer(1, o) : dc_residual(dt, 0, nullptr, 1, 1, o))) return rc;
        HIP_TRY(hipMemcpyAsync(&ctrl_h[2], ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        if (warm_r0_ready && ctrl_h[2].done) {
          // T^n is already converged: its iterate with the BC ghosts in u[1]
          if ((rc = dc_residual(dt, 0, nullptr, 1, 1, o))) return rc;
          HIP_TRY(hipMemcpyAsync(&ctrl_h[2], ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
          HIP_TRY(hipStreamSynchronize(stream));
        }
        bool used = false;
        if (!ctrl_h[2].done && (rc = warm_start_apply(dt, o, &used))) return rc;
        st.warm_start = used ? 1 : 0;
        hipLaunchKernelGGL(k_copy, dim3((g.total + 255) / 256), dim3(256), 0, stream, g.total, (const double*)u[1],
                           u[0]);
        cur = 1;
        reset_and_validate();
        LAUNCH_CHECK("warm start");
      }
      warm_pending = false;
    }
    if (!o.subcycle_conductivity) {
      // iteration-0 alpha and parameters from the BC-applied u (PseudoTransientSolver.jl:254-261)
      for (int t = 0; t < 3; ++t)
        if (!frozen[t]) {
          HIP_TRY(hipMalloc((void**)&frozen[t], sizeof(double) * g.total));
          owned.push_back(frozen[t]);
          hipLaunchKernelGGL(k_fill, dim3((g.total + 255) / 256), dim3(256), 0, stream, g.total, 0.0, frozen[t]);
        }
      hipLaunchKernelGGL(k_halo_fill, grid_full(g), dim3(256), 0, stream, g, gh, u[cur], 0);
      hipLaunchKernelGGL(k_alpha_full, grid_full(g), dim3(256), 0, stream, g, kap, (const double*)rho,
                         (const double*)nullptr, (const double*)u[cur], frozen[0]);
      IterConst ic{L, Vpdtau, dt};
      hipLaunchKernelGGL(k_iter_params_full, grid_full(g), dim3(256), 0, stream, g, ic, (const double*)rho,
                         (const double*)frozen[0], frozen[2], frozen[1]);
      LAUNCH_CHECK("frozen parameters");
    }
    int rc = run_loop64(o, dt, Vpdtau, 0);
    if (rc) return rc;
  } else if (o.precision != CDIFF_PREC_FP32_DC) {
    return fail(CDIFF_BAD_ARG, "unknown precision mode");
  }

  const Ctrl& c = ctrl_h[2];
  const bool dc = o.precision == CDIFF_PREC_FP32_DC && !fallback;
  st.fp64_fallback = fallback ? 1 : 0;
  if (!dc) {
    step_r_start = c.init;
    st.rel_err = c.rel;
    st.abs_err = c.abs_;
    st.niter = c.final_iter;
    st.nchecks = c.nchecks;
    st.outer = 0;
  }
  const int status = c.status;
  if (status != CDIFF_OK) {
    if (stats) *stats = st;
    char buf[256];
    if (status == CDIFF_NONFINITE)
      return fail(status, "Non-finite error detected! abs_error = " + julia_float(c.abs_) +
                              ", rel_error = " + julia_float(c.rel) + ", exiting...");
    if (status == CDIFF_MAX_ITER) {
      std::snprintf(buf, sizeof buf, "Maximum iteration limit reached (%d), ", o.max_iter);
      return fail(status, std::string(buf) + "abs_error = " + julia_float(c.abs_) + ", rel_error = " +
                              julia_float(c.rel) + ", exiting...");
    }
    return validation_message(status);
  }
  const int fi = c.final_iter;
  int src = (int)((cur + fi - 1) & 1), dst = 1 - src;
  const double* pre = nullptr;
  if (!dc) {
    // the launch that ended at the final iteration (checks always end a launch)
    for (const auto& L : launches)
      if (L.last_iter == fi) {
        dst = L.dst;
        src = 1 - dst;
        pre = L.pair ? u_pre : u[src];
      }
    if (!pre) pre = u[src];
  }
  last_iter = fi;
  if (dc) {
    // final iterate (with the BC ghosts of its own domain values) in u[dst];
    // alpha of that state over the whole array; the FP64 damped flux restarts
    // from the plain flux of that state (the APT fixed point of q), so a
    // following FP64 step warm-starts consistently
    dst = dc_final_buf;
    src = 1 - dst;
    double* const qd[3] = {q[0][dst], q[1][dst], ndim == 3 ? q[2][dst] : nullptr};
    launch_dc_end(g, kap, rho, (int)(mean == CDIFF_MEAN_HARMONIC), u[dst], u[src], alpha, qd, stream);
    LAUNCH_CHECK("dc epilogue");
  } else if (o.subcycle_conductivity) {
    hipLaunchKernelGGL(k_alpha_full, grid_full(g), dim3(256), 0, stream, g, kap, (const double*)rho, pre,
                       (const double*)u[dst], alpha);
  } else {
    hipLaunchKernelGGL(k_copy, dim3((g.total + 255) / 256), dim3(256), 0, stream, g.total, (const double*)frozen[0],
                       alpha);
  }
  // T = u (positivity, cutoff!) and the relative change for next_dt in one pass
  hipLaunchKernelGGL(k_step_end, dim3(step_blocks), dim3(STEP_THREADS), 0, stream, g, (const double*)u[dst], Tdev,
                     o.enforce_positivity, o.apply_cutoff, (const double*)u_prev, (const double*)red, o.tol,
                     o.calculate_next_dt ? red + 8 : nullptr);
  LAUNCH_CHECK("epilogue");
  if (o.calculate_next_dt) {
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(256), 0, stream, (const double*)(red + 8), step_blocks, red + 1);
    hipLaunchKernelGGL(k_next_dt_final, dim3(1), dim3(1), 0, stream, (const double*)(red + 1), dt,
                       o.maximum_u_change, o.timestep_growth_factor, red + 2);
    LAUNCH_CHECK("next_dt");
    HIP_TRY(hipMemcpyAsync(red_h, red + 2, sizeof(double), hipMemcpyDeviceToHost, stream));
  }
  HIP_TRY(hipStreamSynchronize(stream));
  if (next_dt_out) *next_dt_out = o.calculate_next_dt ? red_h[0] : dt;
  cur = dst;
  last_src = src;
  last_pre = dc ? u[dst] : pre;
  have_last = true;
  have_prev = true;  // u_prev = T^n, the previous level of the next step
  dt_prev = dt;
  t_written = true;
  warm_account(st, dc);
  const auto t1 = std::chrono::steady_clock::now();
  st.time = std::chrono::duration<double>(t1 - t0).count();
  if (stats) *stats = st;
  return CDIFF_OK;
}

int cdiff_solver::get_field(int field, double* out) {
  const int64_t n = (int64_t)g.n[0] * g.n[1] * g.n[2];
  auto dl = [&](const double* d) -> int {
    HIP_TRY(hipMemcpy2DAsync(out, sizeof(double) * g.n[0], d + g.off, sizeof(double) * g.pitch,
                             sizeof(double) * g.n[0], (size_t)g.n[1] * g.n[2], hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return CDIFF_OK;
  };
  switch (field) {
    case CDIFF_FIELD_U: return dl(Tdev);
    case CDIFF_FIELD_U_PREV: return dl(u_prev);
    case CDIFF_FIELD_ALPHA: return dl(alpha);
    case CDIFF_FIELD_SOURCE:
      if (S) return dl(S);
      std::fill(out, out + n, 0.0);
      return CDIFF_OK;
    case CDIFF_FIELD_DENSITY:
      if (rho) return dl(rho);
      std::fill(out, out + n, 1.0);
      return CDIFF_OK;
    default: break;
  }
  if (field >= CDIFF_FIELD_Q && field < CDIFF_FIELD_Q + ndim) return dl(q[field - CDIFF_FIELD_Q][cur]);
  if (field >= CDIFF_FIELD_CACHE && field < CDIFF_FIELD_CACHE + ndim)
    return dl(coef[(ndim == 2 ? 3 : 6) + field - CDIFF_FIELD_CACHE]);
  const bool want_params = field == CDIFF_FIELD_DTAU_RHO || field == CDIFF_FIELD_THETA_R_DTAU;
  const bool want_qp = field >= CDIFF_FIELD_QPRIME && field < CDIFF_FIELD_QPRIME + ndim;
  const bool want_res = field == CDIFF_FIELD_RES;
  if (!want_params && !want_qp && !want_res) return fail(CDIFF_BAD_ARG, "unknown field");
  // scratch: merged pre-update iterate, its alpha, and the plain fluxes
  std::vector<double*> tmp;
  auto scratch = [&]() -> double* {
    double* p = nullptr;
    if (hipMalloc((void**)&p, sizeof(double) * g.total) != hipSuccess) return nullptr;
    hipLaunchKernelGGL(k_fill, dim3((g.total + 255) / 256), dim3(256), 0, stream, g.total, 0.0, p);
    tmp.push_back(p);
    return p;
  };
  auto release = [&]() {
    (void)hipStreamSynchronize(stream);
    for (double* p : tmp) (void)hipFree(p);
  };
  int rc = CDIFF_OK;
  if (want_params) {
    double* d = scratch();
    double* t = scratch();
    if (!d || !t) {
      release();
      return fail(CDIFF_HIP_ERROR, "hipMalloc(scratch)");
    }
    IterConst ic{L, last_Vpdtau, last_dt};
    hipLaunchKernelGGL(k_iter_params_full, grid_full(g), dim3(256), 0, stream, g, ic, (const double*)rho,
                       (const double*)alpha, d, t);
    rc = dl(field == CDIFF_FIELD_DTAU_RHO ? d : t);
    release();
    return rc;
  }
  if (!have_last) {
    // nothing computed yet: q' and res are still the constructor's zeros
    std::fill(out, out + n, 0.0);
    return CDIFF_OK;
  }
  const int src = last_src, dst = 1 - last_src;
  double* qp[3] = {nullptr, nullptr, nullptr};
  for (int a = 0; a < ndim; ++a) qp[a] = scratch();
  double* a_pre = scratch();
  double* res = scratch();
  for (double* p : tmp)
    if (!p) {
      release();
      return fail(CDIFF_HIP_ERROR, "hipMalloc(scratch)");
    }
  // merged pre-update iterate: domain from u[src], ghosts from u[dst]
  double* merged = scratch();
  if (!merged) {
    release();
    return fail(CDIFF_HIP_ERROR, "hipMalloc(scratch)");
  }
  hipLaunchKernelGGL(k_merge, grid_full(g), dim3(256), 0, stream, g, last_pre ? last_pre : (const double*)u[src],
                     (const double*)u[dst], merged);
  if (last_subcycle)
    hipLaunchKernelGGL(k_alpha_full, grid_full(g), dim3(256), 0, stream, g, kap, (const double*)rho,
                       (const double*)nullptr, (const double*)merged, a_pre);
  else
    hipLaunchKernelGGL(k_copy, dim3((g.total + 255) / 256), dim3(256), 0, stream, g.total, (const double*)alpha,
                       a_pre);
  for (int a = 0; a < ndim; ++a)
    hipLaunchKernelGGL(k_plain_flux, grid_full(g), dim3(256), 0, stream, g, a, (int)(mean == CDIFF_MEAN_HARMONIC),
                       (const double*)merged, (const double*)a_pre, qp[a]);
  if (want_qp) {
    rc = dl(qp[field - CDIFF_FIELD_QPRIME]);
    release();
    return rc;
  }
  Coefs cf{};
  for (int t = 0; t < 9; ++t) cf.c[t] = coef[t];
  hipLaunchKernelGGL(k_residual_field, grid_full(g), dim3(256), 0, stream, g, last_dt, (const double*)u[dst],
                     (const double*)u_prev, (const double*)qp[0], (const double*)qp[1],
                     (const double*)(ndim == 3 ? qp[2] : qp[0]), cf, (const double*)S, (const double*)tzA,
                     (const double*)tzB, (const double*)a_pre, res);
  rc = dl(res);
  release();
  return rc;
}

int cdiff_solver::bench_sweeps(double dt, int precision, int check, int n, double* mean_ms) {
  if (!metrics_set) return fail(CDIFF_NOT_READY, "cdiff_set_metrics has not been called");
  if (n < 1 || !(dt > 0)) return fail(CDIFF_BAD_ARG, "nsweeps >= 1 and dt > 0 required");
  cdiff_opts o;
  cdiff_default_opts(&o);
  double smin = spacing[0];
  for (int a = 1; a < ndim; ++a) smin = std::min(smin, spacing[a]);
  const double Vpdtau = o.cfl * smin;
  hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(1), 0, stream, ctrl, 0.0, 0.0);
  if (ctrl_in) hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(1), 0, stream, ctrl_in, 0.0, 0.0);
  LAUNCH_CHECK("k_ctrl_reset");
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  const bool pair = (check & 2) != 0;
  const bool lin_sweep = (check & 4) != 0 && precision != CDIFF_PREC_FP64;
  if (pair && (!pairs_ok() || (n & 1)))
    return fail(CDIFF_BAD_ARG, "two-iteration launches need a 2D mesh without periodic faces and an even nsweeps");
  if (lin_sweep && (!lin || !lin_coefs_ready || n % lin_k))
    return fail(CDIFF_BAD_ARG, "coefficient-form sweeps need a prior FP32 step and nsweeps a multiple of their depth");
  int rc = CDIFF_OK;
  int nl = 0;
  const int per = lin_sweep ? lin_k : (pair ? 2 : 1);
  auto one = [&]() {
    const int src = (cur + nl) & 1;
    if (lin_sweep)
      rc = launch_lin(dt, nl & 1, (check & 1) != 0, lin_k);
    else if (precision == CDIFF_PREC_FP64)
      rc = launch_sweep64(o, dt, Vpdtau, src, (check & 1) != 0, pair);
    else
      rc = launch_sweep32(dt, Vpdtau, src, (check & 1) != 0, pair);
    ++nl;
  };
  // one untimed launch first: the GPU is busy with it while the timed ones
  // are enqueued, so the events time the launches and not the host
  one();
  (void)hipEventRecord(e0, stream);
  for (int t = 0; t < n && rc == CDIFF_OK; t += per) one();
  (void)hipEventRecord(e1, stream);
  hipError_t se = hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (se != hipSuccess) return hip_fail(se, "hipEventSynchronize");
  cur = (cur + nl) & 1;
  have_last = false;
  *mean_ms = ms / n;
  return CDIFF_OK;
}
