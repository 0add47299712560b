// Dyson-step, predictor, mean-field and bootstrap entry points (FP64).
#pragma once
#include "ctx.h"

namespace ned {

constexpr int KMAX = 8;

// Integrator tables passed by value to kernels.
struct WTab {
  int k;
  double start[(KMAX + 1) * (KMAX + 1)];
  double Omega[(KMAX + 1) * (KMAX + 1)];
  double pdiff[(KMAX + 1) * (KMAX + 1)];
  double bd[KMAX + 2];
  double ex[KMAX + 1];
};

WTab make_wtab(int k);

void set_matsubara(Ctx &c, double beta);
void dyson_step(Ctx &c, int tstp, double mu, double beta, double dt, double *err3);
void extrapolate(Ctx &c, int tstp);
void hf_decomp(Ctx &c, int tstp, const double *h0_host);
double dyson_start(Ctx &c, double mu, double beta, double dt);
void energy(Ctx &c, int tstp, const double *h0_host, double beta, double dt, double *out3);
void invalidate_minv(Ctx &c);
// mark hmf(t) / Sigma slice t as changed: drops the cached inverses that use it
void touch_time(Ctx &c, int t);

}  // namespace ned
