// temporary: 3D sweep and FP32 defect correction are not wired yet
#include "solver.h"
int cdiff_solver::step_fp32_dc(double, const cdiff_opts&, double, cdiff_stats*) {
  return fail(CDIFF_BAD_ARG, "FP32 defect correction not available");
}
int cdiff_solver::launch3d_fp64(const cdiff_opts&, double, double, int, bool) {
  return fail(CDIFF_BAD_ARG, "3D sweep not available");
}
