// The accept / reject / lambda logic of LevenbergMarquardtSolver::Iterate
// (opt/levenberg_marquardt_solver.tcc:239-340) as one host/device function, shared by the device
// optimizer's decision kernel and the batched persistent kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/symhip_types.h"

namespace symhip {

struct LmDecision {
  double lambda, nu;
  symhip_iteration_stats rec;
  int32_t status, reason, accept, set_best;
};

// e0: error at Init, e1: error at New, e_best: error at Best, lde: sum u (rhs - damping u).
// nl2 / nu2 / dotlu: |last update|^2, |update|^2, last.update (bold updates only).
__host__ __device__ inline LmDecision lm_decide(double e0, double e1, double e_best, double lde,
                                                double lambda, double nu, int iteration,
                                                const symhip_optimizer_params& p,
                                                int have_last_update, double nl2, double nu2,
                                                double dotlu, double eps) {
  LmDecision d{};
  const double rel = (e0 - e1) / (e0 + eps);
  const double nel = e0 + 0.5 * lde;
  const double gain = (e0 - e1) / (e0 - nel);
  d.rec.iteration = iteration;
  d.rec.current_lambda = lambda;
  d.rec.new_error = e1;
  d.rec.new_error_linear = nel;
  d.rec.relative_reduction = rel;
  int status = 0, reason = 0;
  if (rel > -p.early_exit_min_reduction / 10 && rel < p.early_exit_min_reduction)
    status = SYMHIP_STATUS_SUCCESS;
  else if (e1 < p.early_exit_min_absolute_error)
    status = SYMHIP_STATUS_SUCCESS;
  bool accept = rel > 0;
  double angle = 0;
  if (p.enable_bold_updates && have_last_update && !accept) {
    angle = (nl2 > 0 && nu2 > 0) ? dotlu / (sqrt(nl2) * sqrt(nu2)) : 0.0;
    accept = ((1 - angle) * (1 - angle) * e1) <= e_best;
  }
  if (!accept && lambda >= p.lambda_upper_bound) {
    status = SYMHIP_STATUS_FAILED;
    reason = SYMHIP_FAILURE_LAMBDA_OUT_OF_BOUNDS;
  }
  int set_best = 0;
  if (!accept) {
    if (p.lambda_update_type == SYMHIP_LAMBDA_STATIC) {
      lambda *= p.lambda_up_factor;
    } else {
      lambda *= nu;
      nu *= 2;
    }
  } else {
    if (p.lambda_update_type == SYMHIP_LAMBDA_STATIC) {
      lambda *= p.lambda_down_factor;
    } else {
      lambda *= fmax(1.0 / p.dynamic_lambda_update_gamma,
                     1.0 - (p.dynamic_lambda_update_beta - 1) *
                               pow(2 * gain - 1, (double)p.dynamic_lambda_update_p));
      nu = 2;
    }
    set_best = e1 <= e_best;
  }
  d.lambda = fmin(fmax(lambda, p.lambda_lower_bound), p.lambda_upper_bound);
  d.nu = nu;
  d.rec.update_angle_change = angle;
  d.rec.update_accepted = accept ? 1 : 0;
  d.status = status;
  d.reason = reason;
  d.accept = accept ? 1 : 0;
  d.set_best = set_best;
  return d;
}

}  // namespace symhip
