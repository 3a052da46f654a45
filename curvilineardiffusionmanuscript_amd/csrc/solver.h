// solver.h -- the libcdiff handle (struct cdiff_solver) and its methods.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/cdiff.h"
#include "cdiff_dev.h"
#include "sweep_args.h"

struct cdiff_solver {
  int device = 0;
  hipStream_t stream = nullptr;
  cdiff::Geo g{};
  int ndim = 2, nhalo = 1, mean = 0;
  double cp = 1.0;
  cdiff_bc bcs[6]{};
  cdiff::Ghost gh{};
  bool dirichlet_outer = false;
  cdiff::KappaLaw kap{};
  double L = 0, spacing[3] = {0, 0, 0};
  bool metrics_set = false;
  double N_full = 0;
  cdiff_tuning tune{};  // kernel variants and defect-correction parameters (cdiff_set_tuning)

  // FP64 state
  double* u[2] = {nullptr, nullptr};
  double* q[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
  double* u_prev = nullptr;
  double* Tdev = nullptr;
  double* alpha = nullptr;
  double* coef[9] = {nullptr};  // 2D: g11 g22 c12 ac bc; 3D: g11 g22 g33 cxe cxz cez ac bc gc
  double *S = nullptr, *rho = nullptr, *tzA = nullptr, *tzB = nullptr;
  double *S_store = nullptr, *rho_store = nullptr, *tzA_store = nullptr, *tzB_store = nullptr;
  double* frozen[3] = {nullptr, nullptr, nullptr};  // alpha, theta, dtau (subcycle_conductivity = false)
  double* part_sum = nullptr;
  double* part_max = nullptr;
  int nparts = 0;
  int part_cap = 0;  // capacity of part_sum / part_max
  double* red = nullptr;
  double* red_h = nullptr;
  cdiff::Ctrl* ctrl = nullptr;
  cdiff::Ctrl* ctrl_h = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};  // [0, 1]: segments read one behind; [2]: a correction round's end
  std::vector<void*> owned;

  // FP32 shadows (defect-correction mode)
  float* u32[2] = {nullptr, nullptr};
  float* q32[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
  float* base32 = nullptr;  // u_prev / Delta-t + S + correction, per cell
  float* coef32[9] = {nullptr};
  float *S32 = nullptr, *rho32 = nullptr, *tzA32 = nullptr, *tzB32 = nullptr;
  float* a32 = nullptr;     // alpha frozen at the FP64 iterate
  float* ub32 = nullptr;    // FP32 copy of the FP64 iterate (Newton correction)
  float* da32 = nullptr;    // d(alpha)/du at the FP64 iterate (Newton correction)
  bool dc_newton = false;
  double* a64 = nullptr;    // alpha of the FP64 iterate
  int64_t dc_inner_guess = 48;  // sweeps of the last correction solve (length of the first segment)
  cdiff::Ghost gh32{};      // ghost rules of the correction
  cdiff::Ctrl* ctrl_in = nullptr;
  double* dc_psum = nullptr;
  double* dc_pmax = nullptr;
  int dc_nparts = 0;
  int dc_final_buf = 0;
  std::vector<void*> owned32;
  bool coefs32_valid = false, aux32_valid = false, q32_valid = false;

  // FP32 correction sweep in coefficient form (2D, no periodic faces)
  bool lin = false;
  int lin_k = 4;  // iterations per launch at most
  int lin_layout = 2;  // 2D sweep: 0, 1 tile layouts (lin2d_tile), 2 marching
  int march_rows = 0, march_chunks = 0, march_strips = 0;  // marching sweep work split
  int lin_tiles = 0, lin_chunks = 0, lin_nblocks = 0;
  float* lin_m[9] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  float* lin_fx[3] = {nullptr, nullptr, nullptr};
  float* lin_fy[3] = {nullptr, nullptr, nullptr};
  float* lin_fz[3] = {nullptr, nullptr, nullptr};
  int lin3_ch = 0, lin3_strips = 0;  // 3D: planes per chunk, 64-wide strips
  // 3D two-iteration sweep: planes per chunk, chunks, tiles, blocks (one block per CU)
  int lin3b_ch = 0, lin3b_chunks = 0, lin3b_tiles = 0, lin3b_nblocks = 0;
  int lin_last_parts = 0;  // partial sums written by the last correction-sweep launch
  float* lin_cc = nullptr;
  bool lin_metrics_valid = false;
  bool lin_coefs_ready = false;
  double* lin_psum = nullptr;
  double* lin_pmax = nullptr;
  int lin_nparts = 0;
  double dc_Vpdtau = 0.0;
  // extra scales of the correction sweep's dtau and theta (the solver's
  // defaults, negative until the first mixed-precision step reads them; 1
  // after the sweeps diverged once)
  double dc_dtau_scale = 1.0, dc_theta_scale = 1.0;
  double dc_dtau_scale_dflt = -1.0, dc_theta_scale_dflt = -1.0;
  double dc_cfl_scale = -1.0;  // pseudo-time step scale of the correction sweeps (< 0: not chosen yet)

  // sweep decomposition
  bool tile2d = true;  // 2D: LDS-tiled kernel (else the row-marching kernel)
  bool col3d = true;   // 3D: 64 x 8 column kernel (else one lane per column with lane/wave rings)
  int ncu = 256;       // compute units of the device
  int strips = 0, chunks = 0, ch = 0, ntiles = 0, nblocks = 0;
  int strips3[2] = {0, 0};

  // persistent solver state
  int cur = 0;        // ping-pong index of the current q (and u) buffers
  int last_iter = 0;  // solver.last_err_check
  bool have_last = false;
  int last_src = 0;  // buffer written by the last launch's source (pre-update iterate, single launches)
  double* u_pre = nullptr;       // pre-update iterate of a two-iteration launch ending in a check
  const double* last_pre = nullptr;  // domain values of the iterate before the last iteration
  double last_dt = 0;
  double last_Vpdtau = 0;
  int last_subcycle = 1;
  int enforce_pos_last = 0;
  bool t_written = false;

  // graph replay of blocks of FP64 2D pseudo-iterations without a check
  // (tune.graphs): the launches of GRAPH_ITERS iterations from each source
  // buffer parity, captured once and replayed while their arguments (this
  // step's iteration constants come from ic_dev) stay the same
  static constexpr int GRAPH_ITERS = 16;
  cdiff::IterConst* ic_dev = nullptr;   // device copy of the step's iteration constants
  cdiff::IterConst* ic_host = nullptr;  // its pinned staging
  bool graph_capture = false;           // launch_sweep64 is being captured
  struct GraphBlock {
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> key;  // the arguments and launch shape of the capture
  };
  GraphBlock gblock[2];  // by the source buffer of the block's first launch
  void graph_reset();

  struct Launch {
    int last_iter;  // last pseudo-iteration of the launch
    int dst;        // buffer it wrote
    bool pair;      // two iterations
  };
  std::vector<Launch> launches;  // of the current step

  // learned warm start (2D FP32_DC steps)
  cdiff_net* net = nullptr;
  double* jac = nullptr;         // cell Jacobian (2D), for the network's features
  double* rz_buf = nullptr;      // cell-centre radius (2D axisymmetric)
  const double* rz_r = nullptr;  // rz_buf when the geometry is axisymmetric, else null
  bool have_prev = false;        // u_prev holds the previous time level T^{n-1}
  double dt_prev = 0.0;
  double* learn_scratch = nullptr;  // >= 1024 doubles: reductions of the feature pass
  double* learn_ts = nullptr;       // device: [max |T^n|, dt max |r(T^n)|] of the current step
  bool warm_pending = false;        // a prediction for this step is in the network's workspace
  bool dc_failed = false;  // the last mixed-precision step gave up (diverged or did not converge)
  bool warm_r0_ready = false;       // the feature pass left T^n's FP64 residual partial sums in place
  int warm_rejects = 0;             // consecutive consultations that did not pay
  int warm_skip = 0;                // steps left before the network is consulted again
  double warm_gain = 0.0;           // decades of residual the prediction used at this step saved
  double warm_r_pred = 0.0;         // residual of that predicted start
  double warm_decades = 1.0;        // decades a consultation is expected to take off the start
  double work_per_decade = 0.0;     // the last step's work (correction-sweep units) per decade of
                                    // residual reduction; 0 before the first step
  double step_r_start = 0.0;        // FP64 residual at the start of the current step
  void warm_back_off();
  void warm_account(const cdiff_stats& st, bool dc);
  bool warm_pays() const;  // expected saving of a consultation above its cost

  std::string err;

  int fail(int code, const std::string& msg);
  bool lin_ok() const;
  // features of a step of size dt from the resident temperature (Tdev = T^n,
  // u_prev = T^{n-1} when have_prev) into X0; the feature scales into ts_dev[0..1]
  int learn_features(double dt, void* X0, double* ts_dev);
  // (Tdev - u_prev) / ts_dev[1] after a step
  int learn_target(const double* ts_dev, float* tgt);
  int warm_start_apply(double dt, const cdiff_opts& o, bool* used);
  int hip_fail(hipError_t e, const char* what);
  void configure_sweep();
  int ensure_parts();  // part_sum / part_max hold nparts entries
  int step(double dt, const cdiff_opts& o, cdiff_stats* stats, double* next_dt);
  int run_loop64(const cdiff_opts& o, double dt, double Vpdtau, int first_iter_base);
  int enqueue_next_dt(const double* unew, const double* uold, double dt, double tol, double muc, double growth);
  int get_field(int field, double* out);
  int validation_message(int status);
  int step_fp32_dc(double dt, const cdiff_opts& o, double Vpdtau, cdiff_stats* stats);
  int launch_sweep64(const cdiff_opts& o, double dt, double Vpdtau, int src, bool check, bool pair = false);
  cdiff::Sweep2DArgs<double> sweep2d_args(const cdiff_opts& o, double dt, double Vpdtau, int src) const;
  bool graph_ok() const;
  int graph_block(const cdiff_opts& o, double dt, double Vpdtau, int src, bool pair);
  int launch_sweep32(double dt, double Vpdtau, int src, bool check, bool pair = false);
  int launch_lin(double dt, int src, bool check, int nit, bool zero_src = false);
  int lin_coefs(double dt, double Vpdtau, int buf);
  bool pairs_ok() const;
  int bench_sweeps(double dt, int precision, int check, int n, double* mean_ms);
  int ensure_fp32();
  int dc_residual(double dt, int srcbuf, const float* e, int dstbuf, int iter, const cdiff_opts& o);
  int dc_finalize_outer(int iter, const cdiff_opts& o);  // norm of the last outer pass's partial sums
  // the fused outer pass on explicit buffers (2D, lin_ok); finalize: update ctrl
  // e_sel: take the correction from u32[e_sel->final_buf] instead of e
  int dc_outer_pass(double dt, const double* u_old, const float* e, double* u_new, const double* uprev, int iter,
                    const cdiff_opts& o, bool finalize, const cdiff::Ctrl* e_sel = nullptr);
};
