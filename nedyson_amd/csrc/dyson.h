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

// Streams, events and the device copy of the members' operands for stepping
// many systems' Dyson recurrences with one launch per link (all on one
// device).
struct DysonGroup {
  int device;
  hipStream_t s = nullptr, sd = nullptr;  // the grouped chain and its bulk updates
  hipEvent_t evm = nullptr, evs[2] = {nullptr, nullptr};
  std::vector<hipEvent_t> evr;  // one per member: its stream -> the group's
  DevBuf args;                  // the members' ChainArg
  void *host_args = nullptr;    // pinned staging of the same
  size_t host_cap = 0;
  explicit DysonGroup(int device);
  ~DysonGroup();
  DysonGroup(const DysonGroup &) = delete;
  DysonGroup &operator=(const DysonGroup &) = delete;
};
// dyson_step for every context of cs at the same tstp (each with its own mu,
// beta, dt; err3 holds 3 per context): the per-system phases on each
// context's stream, and the retarded and lesser recurrences of all systems
// together, one launch per link and per panel update, with the system in
// blockIdx.z.  The same results as dyson_step on each.
void dyson_step_group(DysonGroup &g, const std::vector<Ctx *> &cs, int tstp, const double *mu, const double *beta,
                      const double *dt, double *err3);
void extrapolate(Ctx &c, int tstp);
void hf_decomp(Ctx &c, int tstp, const double *h0_host);
// hf_decomp in two parts: the h0 upload, and the kernels alone (no host
// copies or waits, for use inside a captured graph)
void hf_upload_h0(Ctx &c, const double *h0_host);
void hf_kernels(Ctx &c, int tstp);
// the update norms of the last dyson_step (its err3), read back on the
// context's stream
void dyson_read_err(Ctx &c, double *err3);
double dyson_start(Ctx &c, double mu, double beta, double dt);
void energy(Ctx &c, int tstp, const double *h0_host, double beta, double dt, double *out3);
// Equilibrium GF2 (corr) or HF Matsubara function on the GLL nodes; writes
// G^M (G.mat), Sigma^M (S.mat), the T = 0 slices of G and Sigma and hmf[0].
void matsubara_gf2(Ctx &c, const double *h1, double nelec, double beta, bool corr, bool fix_mu, double *mu_io,
                   int maxiter, double tol, double *fock_out, int *iters_out, double *err_out);
// Scratch of lu_solve_aug: prow (64 doubles), tinv (64 x 64), tmp
// (64 x (D + nrhs)), pval/pidx (one per workgroup of a step: D/32 + 2
// entries), ticket (one counter, zeroed by the call).
struct LuWork {
  double *prow, *tinv, *tmp, *pval;
  int *pidx;
  unsigned *ticket;
};
// Solves Aug[:, :D] X = Aug[:, D:D+nrhs] in place (row-major D x (D+nrhs));
// returns the first zero pivot (1-based) or 0.
int lu_solve_aug(hipStream_t s, double *Aug, int D, int nrhs, int *ipiv, int *info, const LuWork &ws);
void invalidate_minv(Ctx &c);
// mark hmf(t) / Sigma slice t as changed: drops the cached inverses that use it
void touch_time(Ctx &c, int t);

}  // namespace ned
