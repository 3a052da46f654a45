// cdiff.cpp -- host runtime of libcdiff: the C ABI declared in include/cdiff.h.
//
// The handle owns every device buffer (pitched, 128-byte aligned rows), one
// HIP stream, and the convergence control block.  A time step is
//   prologue   T -> u[0], u[1], u_prev; validity flags of alpha and S
//   loop       one fused sweep per pseudo-iteration (+ a norm/decision kernel
//              on error-check iterations), enqueued in segments; the host
//              reads the device "done" flag of segment s while segment s+1
//              runs, so the GPU never waits for a host decision
//   epilogue   alpha of the last iterate, positivity / cutoff, next_dt, T
// which is PseudoTransientScheme.step! (PseudoTransientSolver.jl:180-351).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cdiff.h"
#include "cdiff_dev.h"
#include "kernels.h"
#include "sweep_args.h"
#include "solver.h"

using namespace cdiff;

namespace {

const char* kStatusText[] = {"ok",           "non-finite error", "maximum iteration limit", "invalid diffusivity",
                             "invalid source term", "bad argument",  "HIP error",               "metrics not set"};

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

}  // namespace

int cdiff_solver::fail(int code, const std::string& msg) {
  err = msg;
  return code;
}

int cdiff_solver::hip_fail(hipError_t e, const char* what) {
  err = std::string(what) + ": " + hipGetErrorString(e);
  return CDIFF_HIP_ERROR;
}

#define HIP_TRY(expr)                                  \
  do {                                                 \
    hipError_t _e = (expr);                            \
    if (_e != hipSuccess) return h->hip_fail(_e, #expr); \
  } while (0)

#define LAUNCH_CHECK(name)                                                \
  do {                                                                    \
    hipError_t _e = hipGetLastError();                                    \
    if (_e != hipSuccess) return h->hip_fail(_e, "launch of " name);      \
  } while (0)

static int alloc_field(cdiff_solver* h, double** p) {
  HIP_TRY(hipMalloc((void**)p, sizeof(double) * h->g.total));
  h->owned.push_back(*p);
  hipLaunchKernelGGL(k_fill, dim3((h->g.total + 255) / 256), dim3(256), 0, h->stream, h->g.total, 0.0, *p);
  LAUNCH_CHECK("k_fill");
  return CDIFF_OK;
}

static int upload(cdiff_solver* h, double* dst, const double* src) {
  const Geo& g = h->g;
  HIP_TRY(hipMemcpy2DAsync(dst + g.off, sizeof(double) * g.pitch, src, sizeof(double) * g.n[0],
                           sizeof(double) * g.n[0], (size_t)g.n[1] * g.n[2], hipMemcpyHostToDevice, h->stream));
  return CDIFF_OK;
}

static int download(cdiff_solver* h, double* dst, const double* src) {
  const Geo& g = h->g;
  HIP_TRY(hipMemcpy2DAsync(dst, sizeof(double) * g.n[0], src + g.off, sizeof(double) * g.pitch,
                           sizeof(double) * g.n[0], (size_t)g.n[1] * g.n[2], hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return CDIFF_OK;
}

static dim3 grid_full(const Geo& g) {
  const int64_t n = (int64_t)g.n[0] * g.n[1] * g.n[2];
  return dim3((unsigned)((n + 255) / 256));
}

// Ghost rules from the BC tuple, applied in tuple order (boundary_operators.jl:26-30).
static int build_ghost(cdiff_solver* h) {
  Ghost& gh = h->gh;
  for (int ax = 0; ax < h->ndim; ++ax) {
    const cdiff_bc& bl = h->bcs[2 * ax];
    const cdiff_bc& bh = h->bcs[2 * ax + 1];
    const int lo = h->g.lo[ax], hi = h->g.hi[ax];
    for (int side = 0; side < 2; ++side) {
      const int f = 2 * ax + side;
      const cdiff_bc& b = side ? bh : bl;
      gh.val[f] = b.value;
      if (b.kind < 0 || b.kind > 3) return h->fail(CDIFF_BAD_ARG, "unknown boundary condition kind");
      if (b.kind == CDIFF_BC_FIXED_NEG_SYMMETRY && h->ndim != 2)
        return h->fail(CDIFF_BAD_ARG, "FixedNegSymmetryBC is defined for 2D meshes only");
      if (side == 0) {
        switch (b.kind) {
          case CDIFF_BC_PERIODIC: gh.mode[f] = G_COPY; gh.src[f] = hi; break;
          case CDIFF_BC_NEUMANN: gh.mode[f] = G_COPY; gh.src[f] = lo; break;
          case CDIFF_BC_DIRICHLET: gh.mode[f] = G_CONST; gh.src[f] = lo; break;
          default: gh.mode[f] = G_NEG; gh.src[f] = lo; break;
        }
      } else {
        if (b.kind != CDIFF_BC_PERIODIC) {
          switch (b.kind) {
            case CDIFF_BC_NEUMANN: gh.mode[f] = G_COPY; gh.src[f] = hi; break;
            case CDIFF_BC_DIRICHLET: gh.mode[f] = G_CONST; gh.src[f] = hi; break;
            default: gh.mode[f] = G_NEG; gh.src[f] = hi; break;
          }
        } else if (bl.kind == CDIFF_BC_PERIODIC) {
          gh.mode[f] = G_COPY;
          gh.src[f] = lo;
        } else {
          gh.mode[f] = G_KEEP;
          gh.src[f] = hi;
        }
      }
    }
  }
  h->dirichlet_outer = false;
  if (h->ndim == 2 && h->nhalo > 1)
    for (int f = 0; f < 4; ++f)
      if (h->bcs[f].kind == CDIFF_BC_DIRICHLET) h->dirichlet_outer = true;
  return CDIFF_OK;
}

static void set_law(cdiff_solver* h, const cdiff_kappa* kappa, double cp) {
  KappaLaw& k = h->kap;
  k.kappa0 = kappa->kappa0;
  k.t_exp = kappa->t_exp;
  k.rho_exp = kappa->rho_exp;
  k.cp = cp;
  auto int_exp = [](double p) { return (p == std::floor(p) && p >= 0 && p <= 3) ? (int)p : -1; };
  k.t_int = int_exp(kappa->t_exp);
  k.rho_int = int_exp(kappa->rho_exp);
  // T^(n+1/2) by x^n*sqrt(x) unless the exact pow() is requested
  const double f2 = 2.0 * kappa->t_exp;
  const bool half = f2 == std::floor(f2) && kappa->t_exp != std::floor(kappa->t_exp) && kappa->t_exp > 0 &&
                    kappa->t_exp < 4;
  k.t_half = (half && !h->tune.exact_pow) ? (int)std::floor(kappa->t_exp) : -1;
  h->cp = cp;
}

extern "C" {

int cdiff_set_kappa(cdiff_solver* h, const cdiff_kappa* kappa, double cp) {
  if (!h || !kappa) return CDIFF_BAD_ARG;
  set_law(h, kappa, cp);
  return CDIFF_OK;
}

void cdiff_default_opts(cdiff_opts* o) {
  o->max_iter = 1500;
  o->norm_kind = 0;
  o->rtol = 1e-5;
  o->atol = std::sqrt(2.220446049250313e-16);
  o->cfl = 1.0 / std::sqrt(3.0);
  o->apply_cutoff = 0;
  o->calculate_next_dt = 1;
  o->subcycle_conductivity = 1;
  o->enforce_positivity = 0;
  o->tol = 1e-3;
  o->maximum_u_change = 0.05;
  o->timestep_growth_factor = 1.1;
  o->precision = CDIFF_PREC_FP64;
  o->inner_iters = 0;
}

void cdiff_default_tuning(cdiff_tuning* t) {
  if (!t) return;
  std::memset(t, 0, sizeof *t);
  t->sweep2d = 0;
  t->tile_rows = 8;
  t->march_rows64 = 0;
  t->pairs = 1;
  t->sweep3d = 0;
  t->planes3d = 0;
  t->seg_min = 4;
  t->seg_max = 256;
  t->dc_coefficient_form = 1;
  t->dc_newton = 1;
  t->dc_fused3d = 1;
  t->dc_depth2d = 8;
  t->dc_layout2d = 2;
  t->dc_march_rows = 0;
  t->dc_launch_iters = 0;
  t->dc_depth3d = 2;
  t->dc_planes3d = 0;
  t->dc_planes3d_pair = 0;
  t->dc_outer_planes3d = 0;
  t->dc_check = 0;
  t->dc_max_outer = 100;
  t->verbose = 0;
  t->exact_pow = 0;
  t->dc_eta = 0.0;
  t->dc_floor = 0.3;
  t->dc_cfl_scale = 1.2;
  t->dc_dtau_scale = 1.3;
  t->dc_theta_scale = 1.0;
  t->graphs = 0;
  t->warm_policy = 0;
}

int cdiff_set_tuning(cdiff_solver* h, const cdiff_tuning* t) {
  if (!h || !t) return CDIFF_BAD_ARG;
  const cdiff_tuning& v = *t;
  auto bad = [&](const char* what) { return h->fail(CDIFF_BAD_ARG, std::string("cdiff_set_tuning: ") + what); };
  if (v.sweep2d < 0 || v.sweep2d > 1) return bad("sweep2d must be 0 or 1");
  if (v.tile_rows != 4 && v.tile_rows != 8 && v.tile_rows != 16) return bad("tile_rows must be 4, 8 or 16");
  if (v.sweep3d < 0 || v.sweep3d > 1) return bad("sweep3d must be 0 or 1");
  if (v.march_rows64 < 0 || v.planes3d < 0 || v.dc_march_rows < 0 || v.dc_planes3d < 0 ||
      v.dc_planes3d_pair < 0 || v.dc_outer_planes3d < 0)
    return bad("chunk sizes must be >= 0");
  if (v.seg_min < 1 || v.seg_max < v.seg_min) return bad("need 1 <= seg_min <= seg_max");
  if (v.dc_depth2d != 2 && v.dc_depth2d != 4 && v.dc_depth2d != 6 && v.dc_depth2d != 8)
    return bad("dc_depth2d must be 2, 4, 6 or 8");
  if (v.dc_layout2d < 0 || v.dc_layout2d > 2) return bad("dc_layout2d must be 0, 1 or 2");
  if (v.dc_depth3d < 1 || v.dc_depth3d > 2) return bad("dc_depth3d must be 1 or 2");
  if (v.dc_launch_iters < 0 || v.dc_check < 0 || v.dc_max_outer < 1) return bad("dc counts out of range");
  if (v.warm_policy < 0 || v.warm_policy > 1) return bad("warm_policy must be 0 or 1");
  if (!(v.dc_eta >= 0.0 && v.dc_eta < 1.0) || !(v.dc_floor >= 0.0) || !(v.dc_cfl_scale > 0.0) ||
      !(v.dc_dtau_scale > 0.0) || !(v.dc_theta_scale > 0.0))
    return bad("dc_eta in [0, 1), dc_floor >= 0 and positive scales required");
  const bool pow_changed = (v.exact_pow != 0) != (h->tune.exact_pow != 0);
  h->tune = v;
  if (pow_changed) {
    cdiff_kappa k{h->kap.kappa0, h->kap.t_exp, h->kap.rho_exp};
    set_law(h, &k, h->cp);
  }
  // the FP64 launch geometry and the correction parameters are chosen again
  h->graph_reset();
  h->configure_sweep();
  int rc = h->ensure_parts();
  if (rc) return rc;
  h->dc_cfl_scale = -1.0;
  h->dc_dtau_scale_dflt = h->dc_theta_scale_dflt = -1.0;
  return CDIFF_OK;
}

int cdiff_get_tuning(const cdiff_solver* h, cdiff_tuning* t) {
  if (!h || !t) return CDIFF_BAD_ARG;
  *t = h->tune;
  return CDIFF_OK;
}

const char* cdiff_status_string(int s) {
  if (s < 0 || s > 7) return "unknown status";
  return kStatusText[s];
}

const char* cdiff_last_error(const cdiff_solver* h) { return h ? h->err.c_str() : "null handle"; }

int cdiff_create(int ndim, const int64_t* dims_full, int nhalo, int mean, const cdiff_bc* bcs,
                 const cdiff_kappa* kappa, double cp, cdiff_solver** out) {
  if (!out) return CDIFF_BAD_ARG;
  *out = nullptr;
  if ((ndim != 2 && ndim != 3) || !dims_full || !bcs || !kappa || nhalo < 1) return CDIFF_BAD_ARG;
  for (int a = 0; a < ndim; ++a)
    if (dims_full[a] < 2 * nhalo + 1 || dims_full[a] > (1 << 20)) return CDIFF_BAD_ARG;
  if (mean != CDIFF_MEAN_HARMONIC && mean != CDIFF_MEAN_ARITHMETIC) return CDIFF_BAD_ARG;
  cdiff_solver* h = new cdiff_solver();
  HIP_TRY(hipGetDevice(&h->device));
  {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device) == hipSuccess && ncu > 0)
      h->ncu = ncu;
  }
  HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  h->ndim = ndim;
  h->nhalo = nhalo;
  h->mean = mean;
  h->cp = cp;
  for (int f = 0; f < 2 * ndim; ++f) h->bcs[f] = bcs[f];
  Geo& g = h->g;
  g.ndim = ndim;
  for (int a = 0; a < 3; ++a) {
    g.n[a] = a < ndim ? (int)dims_full[a] : 1;
    g.lo[a] = a < ndim ? nhalo : 0;
    g.hi[a] = a < ndim ? g.n[a] - nhalo - 1 : 0;
  }
  g.pitch = round_up(g.n[0], ALIGN);
  g.plane = g.pitch * g.n[1];
  g.off = ALIGN - nhalo;
  g.total = round_up(g.off + g.plane * g.n[2], ALIGN) + ALIGN;
  if (g.total >= ((int64_t)1 << 31)) {
    *out = h;
    return h->fail(CDIFF_BAD_ARG, "mesh too large: a field would hold 2^31 or more elements");
  }
  h->N_full = (double)g.n[0] * g.n[1] * g.n[2];
  cdiff_default_tuning(&h->tune);
  set_law(h, kappa, cp);
  int rc = build_ghost(h);
  if (rc) {
    *out = h;
    return rc;
  }
  for (int b = 0; b < 2; ++b)
    if ((rc = alloc_field(h, &h->u[b]))) return rc;
  for (int a = 0; a < ndim; ++a)
    for (int b = 0; b < 2; ++b)
      if ((rc = alloc_field(h, &h->q[a][b]))) return rc;
  if ((rc = alloc_field(h, &h->u_prev))) return rc;
  if ((rc = alloc_field(h, &h->Tdev))) return rc;
  if ((rc = alloc_field(h, &h->alpha))) return rc;
  // the divergence coefficients in one block, array c at coef[0] + c * total
  const int ncoef = ndim == 2 ? 5 : 9;
  {
    double* blk = nullptr;
    HIP_TRY(hipMalloc((void**)&blk, sizeof(double) * h->g.total * ncoef));
    h->owned.push_back(blk);
    hipLaunchKernelGGL(k_fill, dim3((h->g.total * ncoef + 255) / 256), dim3(256), 0, h->stream, h->g.total * ncoef, 0.0,
                       blk);
    LAUNCH_CHECK("k_fill");
    for (int c = 0; c < ncoef; ++c) h->coef[c] = blk + h->g.total * c;
  }
  HIP_TRY(hipMalloc((void**)&h->ctrl, sizeof(Ctrl)));
  h->owned.push_back(h->ctrl);
  HIP_TRY(hipHostMalloc((void**)&h->ctrl_h, 4 * sizeof(Ctrl), hipHostMallocDefault));
  HIP_TRY(hipMalloc((void**)&h->red, sizeof(double) * (8 + STEP_BLOCKS)));
  h->owned.push_back(h->red);
  HIP_TRY(hipHostMalloc((void**)&h->red_h, sizeof(double) * 8, hipHostMallocDefault));
  for (int e = 0; e < 3; ++e) HIP_TRY(hipEventCreateWithFlags(&h->ev[e], hipEventDisableTiming));
  h->configure_sweep();
  if ((rc = h->ensure_parts())) return rc;
  HIP_TRY(hipMalloc((void**)&h->ic_dev, sizeof(IterConst)));
  h->owned.push_back(h->ic_dev);
  HIP_TRY(hipHostMalloc((void**)&h->ic_host, sizeof(IterConst), hipHostMallocDefault));
  hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(1), 0, h->stream, h->ctrl, 0.0, 0.0);
  LAUNCH_CHECK("k_ctrl_reset");
  HIP_TRY(hipStreamSynchronize(h->stream));
  *out = h;
  return CDIFF_OK;
}

int cdiff_destroy(cdiff_solver* h) {
  if (!h) return CDIFF_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  h->graph_reset();
  for (void* p : h->owned) (void)hipFree(p);
  if (h->ic_host) (void)hipHostFree(h->ic_host);
  for (void* p : h->owned32) (void)hipFree(p);
  if (h->ctrl_h) (void)hipHostFree(h->ctrl_h);
  if (h->red_h) (void)hipHostFree(h->red_h);
  for (int e = 0; e < 3; ++e)
    if (h->ev[e]) (void)hipEventDestroy(h->ev[e]);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return CDIFF_OK;
}

// divergence coefficients (and the 2D cell Jacobian) from metric arrays already
// on the device; frees `tmp` (the metric arrays) when done
static int finish_metrics(cdiff_solver* h, const MetricIn& m, std::vector<double*>& tmp, double L,
                          const double* spacing) {
  const int nd = h->ndim;
  auto cleanup = [&]() {
    for (double* p : tmp) (void)hipFree(p);
    tmp.clear();
  };
  if (nd == 2) {
    // cell Jacobian 1 / det(d(xi, eta)/d(x, y)), a feature of the learned warm start
    if (!h->jac) {
      int rc = alloc_field(h, &h->jac);
      if (rc) {
        cleanup();
        return rc;
      }
    }
    hipLaunchKernelGGL(k_cell_jacobian2d, grid_full(h->g), dim3(256), 0, h->stream, h->g, (const double*)m.cell[0],
                       (const double*)m.cell[1], (const double*)m.cell[2], (const double*)m.cell[3], h->jac);
  }
  double* c[9] = {nullptr};
  for (int t = 0; t < 9; ++t) c[t] = t < (nd == 2 ? 5 : 9) ? h->coef[t] : h->coef[0];
  hipLaunchKernelGGL(k_metric_coefs, grid_full(h->g), dim3(256), 0, h->stream, h->g, m, c[0], c[1], c[2], c[3], c[4],
                     c[5], c[6], c[7], c[8], (const double*)h->rz_r);
  hipError_t le = hipGetLastError();
  hipError_t se = hipStreamSynchronize(h->stream);
  cleanup();
  if (le != hipSuccess) return h->hip_fail(le, "launch of k_metric_coefs");
  if (se != hipSuccess) return h->hip_fail(se, "k_metric_coefs");
  h->L = L;
  for (int a = 0; a < nd; ++a) h->spacing[a] = spacing[a];
  h->metrics_set = true;
  h->coefs32_valid = false;
  h->lin_metrics_valid = false;
  return CDIFF_OK;
}

int cdiff_set_metrics(cdiff_solver* h, const double* const* cell, const double* const* edge, double L,
                      const double* spacing) {
  if (!h || !cell || !edge || !spacing) return CDIFF_BAD_ARG;
  const int nd = h->ndim, nm = nd * nd;
  const int nedge = nd * (nm + 1);
  std::vector<double*> tmp;
  MetricIn m{};
  auto cleanup = [&]() {
    for (double* p : tmp) (void)hipFree(p);
  };
  for (int t = 0; t < nm + nedge; ++t) {
    const double* src = t < nm ? cell[t] : edge[t - nm];
    if (!src) {
      cleanup();
      return h->fail(CDIFF_BAD_ARG, "null metric array");
    }
    double* d = nullptr;
    hipError_t e = hipMalloc((void**)&d, sizeof(double) * h->g.total);
    if (e != hipSuccess) {
      cleanup();
      return h->hip_fail(e, "hipMalloc(metrics)");
    }
    tmp.push_back(d);
    int rc = upload(h, d, src);
    if (rc) {
      cleanup();
      return rc;
    }
    if (t < nm)
      m.cell[t] = d;
    else
      m.edge[t - nm] = d;
  }
  return finish_metrics(h, m, tmp, L, spacing);
}

int cdiff_set_axisymmetric(cdiff_solver* h, const double* r) {
  if (!h) return CDIFF_BAD_ARG;
  if (h->ndim != 2) return h->fail(CDIFF_BAD_ARG, "axisymmetric geometry is 2D (x = r, y = z)");
  if (h->metrics_set) return h->fail(CDIFF_BAD_ARG, "cdiff_set_axisymmetric must precede the metrics");
  if (!r) {
    h->rz_r = nullptr;  // buffer stays allocated in `owned` until destroy
    return CDIFF_OK;
  }
  const Geo& g = h->g;
  for (int j = g.lo[1]; j <= g.hi[1]; ++j)
    for (int i = g.lo[0]; i <= g.hi[0]; ++i)
      if (!(r[i + (int64_t)g.n[0] * j] > 0.0))
        return h->fail(CDIFF_BAD_ARG, "axisymmetric mesh needs r > 0 at every domain cell centre");
  if (!h->rz_buf) {
    int rc = alloc_field(h, &h->rz_buf);
    if (rc) return rc;
  }
  int rc = upload(h, h->rz_buf, r);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->rz_r = h->rz_buf;
  return CDIFF_OK;
}

int cdiff_set_nodes(cdiff_solver* h, const double* const* nodes, double L, const double* spacing) {
  if (!h || !nodes || !spacing) return CDIFF_BAD_ARG;
  const int nd = h->ndim, nm = nd * nd;
  const int nedge = nd * (nm + 1);
  for (int c = 0; c < nd; ++c)
    if (!nodes[c]) return h->fail(CDIFF_BAD_ARG, "null node array");
  const Geo& g = h->g;
  const int64_t s1 = g.n[0] + 1, s2 = s1 * (g.n[1] + 1);
  const int64_t nnodes = s2 * (nd == 3 ? g.n[2] + 1 : 1);
  std::vector<double*> tmp;
  auto cleanup = [&]() {
    for (double* p : tmp) (void)hipFree(p);
  };
  auto dev = [&](size_t n, double** p) -> int {
    hipError_t e = hipMalloc((void**)p, sizeof(double) * n);
    if (e != hipSuccess) {
      cleanup();
      return h->hip_fail(e, "hipMalloc(nodes/metrics)");
    }
    tmp.push_back(*p);
    return CDIFF_OK;
  };
  double* X[3] = {nullptr, nullptr, nullptr};
  int rc;
  for (int c = 0; c < nd; ++c) {
    if ((rc = dev((size_t)nnodes, &X[c]))) return rc;
    hipError_t e = hipMemcpyAsync(X[c], nodes[c], sizeof(double) * nnodes, hipMemcpyHostToDevice, h->stream);
    if (e != hipSuccess) {
      cleanup();
      return h->hip_fail(e, "hipMemcpyAsync(nodes)");
    }
  }
  MetricIn m{};
  MetricOut o{};
  for (int t = 0; t < nm + nedge; ++t) {
    double* d = nullptr;
    if ((rc = dev((size_t)g.total, &d))) return rc;
    if (t < nm) {
      m.cell[t] = o.cell[t] = d;
    } else {
      m.edge[t - nm] = o.edge[t - nm] = d;
    }
  }
  const int64_t ncell = (int64_t)g.n[0] * g.n[1] * (nd == 3 ? g.n[2] : 1);
  if (nd == 2)
    hipLaunchKernelGGL(k_node_metrics2d, dim3((ncell + 255) / 256), dim3(256), 0, h->stream, g, (const double*)X[0],
                       (const double*)X[1], s1, o);
  else
    hipLaunchKernelGGL(k_node_metrics3d, dim3((ncell + 255) / 256), dim3(256), 0, h->stream, g, (const double*)X[0],
                       (const double*)X[1], (const double*)X[2], s1, s2, o);
  hipError_t le = hipGetLastError();
  if (le != hipSuccess) {
    cleanup();
    return h->hip_fail(le, "launch of k_node_metrics");
  }
  return finish_metrics(h, m, tmp, L, spacing);
}

static int set_optional(cdiff_solver* h, double** slot, const double* src) {
  if (!src) {
    *slot = nullptr;  // buffer stays allocated in `owned` until destroy
    return CDIFF_OK;
  }
  if (!*slot) {
    int rc = alloc_field(h, slot);
    if (rc) return rc;
  }
  int rc = upload(h, *slot, src);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  return CDIFF_OK;
}

int cdiff_set_density(cdiff_solver* h, const double* rho) {
  if (!h) return CDIFF_BAD_ARG;
  if (rho) {
    const int64_t n = (int64_t)h->g.n[0] * h->g.n[1] * h->g.n[2];
    bool ones = true;
    for (int64_t e = 0; e < n && ones; ++e) ones = rho[e] == 1.0;
    if (ones) rho = nullptr;  // exactly 1 everywhere: the no-density fast path
  }
  if (!rho) {
    h->rho = nullptr;
    return CDIFF_OK;
  }
  if (!h->rho_store) {
    int rc = alloc_field(h, &h->rho_store);
    if (rc) return rc;
  }
  h->rho = h->rho_store;
  h->aux32_valid = false;
  return set_optional(h, &h->rho, rho);
}

int cdiff_set_source(cdiff_solver* h, const double* S) {
  if (!h) return CDIFF_BAD_ARG;
  if (S) {
    const int64_t n = (int64_t)h->g.n[0] * h->g.n[1] * h->g.n[2];
    bool zeros = true;
    for (int64_t e = 0; e < n && zeros; ++e) zeros = S[e] == 0.0 && !std::signbit(S[e]);
    if (zeros) S = nullptr;
  }
  h->aux32_valid = false;
  if (!S) {
    h->S = nullptr;
    return CDIFF_OK;
  }
  if (!h->S_store) {
    int rc = alloc_field(h, &h->S_store);
    if (rc) return rc;
  }
  h->S = h->S_store;
  return set_optional(h, &h->S, S);
}

int cdiff_set_forcing(cdiff_solver* h, const double* A, const double* B) {
  if (!h) return CDIFF_BAD_ARG;
  h->aux32_valid = false;
  if (!A || !B) {
    h->tzA = h->tzB = nullptr;
    return CDIFF_OK;
  }
  if (!h->tzA_store) {
    int rc = alloc_field(h, &h->tzA_store);
    if (rc) return rc;
    rc = alloc_field(h, &h->tzB_store);
    if (rc) return rc;
  }
  h->tzA = h->tzA_store;
  h->tzB = h->tzB_store;
  int rc = set_optional(h, &h->tzA, A);
  if (rc) return rc;
  return set_optional(h, &h->tzB, B);
}

int cdiff_set_temperature(cdiff_solver* h, const double* T) {
  if (!h || !T) return CDIFF_BAD_ARG;
  h->have_prev = false;
  int rc = upload(h, h->Tdev, T);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  return CDIFF_OK;
}

int cdiff_get_temperature(cdiff_solver* h, double* T) {
  if (!h || !T) return CDIFF_BAD_ARG;
  return download(h, T, h->Tdev);
}

int cdiff_update_conductivity(cdiff_solver* h, const double* T) {
  if (!h || !T) return CDIFF_BAD_ARG;
  int rc = upload(h, h->Tdev, T);
  if (rc) return rc;
  hipLaunchKernelGGL(k_alpha_full, grid_full(h->g), dim3(256), 0, h->stream, h->g, h->kap, h->rho,
                     (const double*)nullptr, h->Tdev, h->alpha);
  LAUNCH_CHECK("k_alpha_full");
  HIP_TRY(hipStreamSynchronize(h->stream));
  return CDIFF_OK;
}

int cdiff_apply_bcs(cdiff_solver* h, double* u) {
  if (!h || !u) return CDIFF_BAD_ARG;
  double* buf = h->u[0];
  int rc = upload(h, buf, u);
  if (rc) return rc;
  hipLaunchKernelGGL(k_halo_fill, grid_full(h->g), dim3(256), 0, h->stream, h->g, h->gh, buf, 1);
  hipLaunchKernelGGL(k_halo_fill, grid_full(h->g), dim3(256), 0, h->stream, h->g, h->gh, buf, 0);
  LAUNCH_CHECK("k_halo_fill");
  return download(h, u, buf);
}

int cdiff_next_dt(cdiff_solver* h, const double* u_new, const double* u_prev, double dt, double tol, double muc,
                  double growth, double* out) {
  if (!h || !u_new || !u_prev || !out) return CDIFF_BAD_ARG;
  int rc = upload(h, h->u[0], u_new);
  if (rc) return rc;
  if ((rc = upload(h, h->u[1], u_prev))) return rc;
  if ((rc = h->enqueue_next_dt(h->u[0], h->u[1], dt, tol, muc, growth))) return rc;
  HIP_TRY(hipMemcpyAsync(h->red_h, h->red + 2, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  *out = h->red_h[0];
  return CDIFF_OK;
}

int cdiff_get_last_iter(cdiff_solver* h, int32_t* out) {
  if (!h || !out) return CDIFF_BAD_ARG;
  *out = h->last_iter;
  return CDIFF_OK;
}

int cdiff_set_last_iter(cdiff_solver* h, int32_t v) {
  if (!h || v < 0) return CDIFF_BAD_ARG;
  h->last_iter = v;
  return CDIFF_OK;
}

int cdiff_reset_state(cdiff_solver* h) {
  if (!h) return CDIFF_BAD_ARG;
  for (int a = 0; a < h->ndim; ++a)
    for (int b = 0; b < 2; ++b) {
      hipLaunchKernelGGL(k_fill, dim3((h->g.total + 255) / 256), dim3(256), 0, h->stream, h->g.total, 0.0, h->q[a][b]);
      LAUNCH_CHECK("k_fill");
    }
  h->last_iter = 0;
  h->cur = 0;
  h->have_last = false;
  h->have_prev = false;
  h->q32_valid = false;
  HIP_TRY(hipStreamSynchronize(h->stream));
  return CDIFF_OK;
}

int cdiff_set_field(cdiff_solver* h, int field, const double* in) {
  if (!h || !in) return CDIFF_BAD_ARG;
  if (field == CDIFF_FIELD_ALPHA) {
    int rc = upload(h, h->alpha, in);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return CDIFF_OK;
  }
  if (field >= CDIFF_FIELD_Q && field < CDIFF_FIELD_Q + h->ndim) {
    int rc = upload(h, h->q[field - CDIFF_FIELD_Q][h->cur], in);
    if (rc) return rc;
    h->q32_valid = false;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return CDIFF_OK;
  }
  return h->fail(CDIFF_BAD_ARG, "field cannot be set");
}

int cdiff_bench_sweeps(cdiff_solver* h, double dt, int precision, int check, int nsweeps, double* mean_ms) {
  if (!h || !mean_ms) return CDIFF_BAD_ARG;
  return h->bench_sweeps(dt, precision, check, nsweeps, mean_ms);
}

int cdiff_get_field(cdiff_solver* h, int field, double* out) {
  if (!h || !out) return CDIFF_BAD_ARG;
  return h->get_field(field, out);
}

int cdiff_step(cdiff_solver* h, double* T, double dt, const cdiff_opts* opts, cdiff_stats* stats, double* next_dt) {
  if (!h || !T) return CDIFF_BAD_ARG;
  int rc = cdiff_set_temperature(h, T);
  if (rc) return rc;
  rc = cdiff_step_resident(h, dt, opts, stats, next_dt);
  if (rc != CDIFF_OK && !h->t_written) return rc;
  int rc2 = cdiff_get_temperature(h, T);
  return rc ? rc : rc2;
}

int cdiff_step_resident(cdiff_solver* h, double dt, const cdiff_opts* opts, cdiff_stats* stats, double* next_dt) {
  if (!h) return CDIFF_BAD_ARG;
  cdiff_opts o;
  if (opts)
    o = *opts;
  else
    cdiff_default_opts(&o);
  return h->step(dt, o, stats, next_dt);
}

}  // extern "C"
