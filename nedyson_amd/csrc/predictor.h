// Timestep driver with the polynomial or learned initial guess (predictor.hip).
#pragma once
#include "mlp.h"
#include "dyson.h"

namespace ned {

int predictor_features();  // network inputs per orbital pair (768)
int predictor_outputs();   // network outputs per orbital pair (128)
void set_predictor(Ctx &c, Mlp *m);
// device pointer to the nao^2 x 128 extrapolation errors of step tstp, or
// nullptr if that step is not among the last seven recorded
const float *predictor_residual(Ctx &c, int tstp);
void do_tstp(Ctx &c, int T, const double *h0, double mu, double beta, double dt, Prec prec, int corr_max,
             double corr_tol, int *iters_out, double *err_out);
// do_tstp for every context of cs at step T (polynomial initial guess), the
// Dyson recurrences of all of them grouped (dyson_step_group): the passes of
// {HF, Sigma} are issued for every system on its own stream, then one grouped
// Dyson step.  corr_tol > 0: a system that converged leaves the group.  A
// system that cannot take step T is left out with NED_ERR_ARG in status, and
// one whose update turns non-finite is dropped with NED_ERR_NONFINITE (the
// others go on).  iters, err, status: one per context.
void batch_do_tstp(DysonGroup &g, const std::vector<Ctx *> &cs, int T, const double *const *h0, const double *mu,
                   const double *beta, const double *dt, Prec prec, int corr_max, double corr_tol, int *iters,
                   double *err, int *status);

}  // namespace ned
