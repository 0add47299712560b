// Timestep driver with the polynomial or learned initial guess (predictor.hip).
#pragma once
#include "mlp.h"

namespace ned {

int predictor_features();  // network inputs per orbital pair (768)
int predictor_outputs();   // network outputs per orbital pair (128)
void set_predictor(Ctx &c, Mlp *m);
// device pointer to the nao^2 x 128 extrapolation errors of step tstp, or
// nullptr if that step is not among the last seven recorded
const float *predictor_residual(Ctx &c, int tstp);
void do_tstp(Ctx &c, int T, const double *h0, double mu, double beta, double dt, Prec prec, int corr_max,
             double corr_tol, int *iters_out, double *err_out);

}  // namespace ned
