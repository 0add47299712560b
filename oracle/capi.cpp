// C entry points of the CPU oracle, for the tests and the CPU-baseline leg of
// the benchmark (loaded with ctypes).  Test infrastructure only.
#include "oracle.h"
#include <cstring>
#include <cmath>

using orc::cplx;

extern "C" {

struct orc_green {
  int nt, ntau, nao, sig;
  cplx *ret, *les, *tv, *mat;
};

static orc::Green view(const orc_green *g) {
  orc::Green v;
  v.nt = g->nt; v.ntau = g->ntau; v.nao = g->nao; v.sig = g->sig;
  v.ret = g->ret; v.les = g->les; v.tv = g->tv; v.mat = g->mat;
  return v;
}

int orc_integrator(int k, double *interp, double *diff, double *integ, double *bd, double *start, double *Omega,
                   double *ex) {
  orc::Integrator I(k);
  int k1 = k + 1;
  std::memcpy(interp, I.poly_interp_.data(), sizeof(double) * k1 * k1);
  std::memcpy(diff, I.poly_diff_.data(), sizeof(double) * k1 * k1);
  std::memcpy(integ, I.poly_integ_.data(), sizeof(double) * k1 * k1 * k1);
  std::memcpy(bd, I.bd_.data(), sizeof(double) * (k + 2));
  std::memcpy(start, I.start_.data(), sizeof(double) * k1 * k1);
  std::memcpy(Omega, I.Omega_.data(), sizeof(double) * k1 * k1);
  std::memcpy(ex, I.ex_.data(), sizeof(double) * k1);
  return 0;
}

void orc_gregory_table(int k, int nmax, double *w) {
  // w[n*(nmax+1)+j] = gregory_weights(n, j) for 0 <= j <= max(n,k), n <= nmax
  orc::Integrator I(k);
  for (int n = 0; n <= nmax; n++)
    for (int j = 0; j <= nmax; j++) w[(size_t)n * (nmax + 1) + j] = (j <= (n > k ? n : k)) ? I.gregory_weights(n, j) : 0.0;
}

void orc_gll(int ntau, double *x, double *mass) {
  orc::Legendre L(ntau);
  std::memcpy(x, L.x.data(), sizeof(double) * (ntau + 1));
  if (mass) std::memcpy(mass, L.mass.data(), sizeof(double) * (ntau + 1) * (ntau + 1));
}

void orc_mixing_kernel(int ntau, int nao, const cplx *B, double beta, int sig, cplx *K) {
  orc::Legendre L(ntau);
  L.mixing_kernel(nao, B, beta, sig, K);
}

void orc_sigma_solve(int nao, int naux, const double *Vija, int tstp, const orc_green *S, const orc_green *G,
                     int nthreads) {
  orc::GF2Decomp s(nao, naux, Vija);
  s.solve(tstp, view(S), view(G), nthreads);
}

void orc_sigma_solve_loop(int nao, int naux, const double *Vija, int tstp, const orc_green *S, const orc_green *G) {
  orc::GF2Decomp s(nao, naux, Vija);
  s.solve_loop(tstp, view(S), view(G));
}

void orc_sigma_point(int nao, int naux, const double *Vija, int comp, int tstp, int t, const orc_green *S,
                     const orc_green *G) {
  orc::GF2Decomp s(nao, naux, Vija);
  s.solve_point(comp, tstp, t, view(S), view(G));
}

void orc_hf(int nao, int naux, const double *Vija, int tstp, cplx *hmf, const cplx *rho, int loop) {
  orc::GF2Decomp s(nao, naux, Vija);
  if (loop) s.solve_HF_loop(tstp, hmf, rho);
  else s.solve_HF(tstp, hmf, rho);
}

void *orc_dyson_new(int nt, int ntau, int nao, int k) { return new orc::Dyson(nt, ntau, nao, k); }
void orc_dyson_free(void *d) { delete (orc::Dyson *)d; }

double orc_dyson_step(void *d, int n, const orc_green *G, const orc_green *S, const cplx *hmf, double mu,
                      double beta, double dt, double *errs) {
  auto *D = (orc::Dyson *)d;
  double e0 = D->dyson_step_ret(n, view(G), view(S), hmf, mu, dt);
  double e1 = D->dyson_step_tv(n, view(G), view(S), hmf, mu, beta, dt);
  double e2 = D->dyson_step_les(n, view(G), view(S), hmf, mu, beta, dt);
  if (errs) { errs[0] = e0; errs[1] = e1; errs[2] = e2; }
  return e0 + e1 + e2;
}

double orc_dyson_start(void *d, const orc_green *G, const orc_green *S, const cplx *hmf, double mu, double beta,
                       double dt) {
  return ((orc::Dyson *)d)->dyson_start(view(G), view(S), hmf, mu, beta, dt);
}

void orc_extrapolate(void *d, int n, const orc_green *G) { ((orc::Dyson *)d)->Extrapolate(n, view(G)); }

double orc_energy_conv(void *d, int tstp, const orc_green *S, const orc_green *G, double beta, double dt) {
  return ((orc::Dyson *)d)->energy_conv(tstp, view(S), view(G), beta, dt);
}

void orc_ctv1(void *d, cplx *ctv, const orc_green *A, const orc_green *Acc, const orc_green *B, int n, double dt) {
  ((orc::Dyson *)d)->CTV1(ctv, view(A), view(Acc), view(B), n, dt);
}

void orc_ctv_tstp(void *d, int tstp, const orc_green *C, const orc_green *A, const orc_green *Acc, const orc_green *B,
                  const orc_green *Bcc, double beta, double dt) {
  ((orc::Dyson *)d)->Ctv_tstp(tstp, view(C), view(A), view(Acc), view(B), view(Bcc), beta, dt);
}

void orc_cles2(void *d, const orc_green *A, const orc_green *Acc, const orc_green *B, const orc_green *Bcc, int m,
               double dt, cplx *res) {
  auto *D = (orc::Dyson *)d;
  int num = m >= D->k ? m : D->k;
  D->Cles2_tstp(0, num, view(A), view(Acc), view(B), view(Bcc), m, dt, res);
}

void orc_cles3(void *d, const orc_green *A, const orc_green *Acc, const orc_green *B, const orc_green *Bcc, int m,
               double beta, cplx *res) {
  auto *D = (orc::Dyson *)d;
  int top = m <= D->k ? D->k : m;
  D->Cles3_tstp(0, top, view(A), view(Acc), view(B), view(Bcc), m, beta, res);
}

void orc_conv_mixing(void *d, cplx *C, const cplx *A, const orc_green *B, double beta) {
  ((orc::Dyson *)d)->conv_mixing(C, A, view(B), beta);
}

void orc_g0_from_h0(void *d, const orc_green *G, double mu, const double *H0, double beta, double h) {
  ((orc::Dyson *)d)->G0_from_h0(view(G), mu, H0, beta, h);
}

void orc_g0_matsubara(void *d, const orc_green *G, double mu, const double *H0, double beta) {
  ((orc::Dyson *)d)->G0_matsubara(view(G), mu, H0, beta);
}

// DecompSimulation::do_boot (src/molNEsim_decomp_impl.h:63-82).  Returns the
// number of iterations used, or -1 when the bootstrap did not converge.
int orc_do_boot(void *d, int naux, const double *Vija, const orc_green *G, const orc_green *S, cplx *hmf,
                const double *fock, double mu, double beta, double dt, int maxiter, double tol, double *err_out,
                int nthreads) {
  auto *D = (orc::Dyson *)d;
  const int nao = D->nao, es = nao * nao, k = D->k;
  orc::GF2Decomp s(nao, naux, Vija);
  double err = 0;
  for (int iter = 0; iter <= maxiter; iter++) {
    for (int t = 0; t <= k; t++) {
      for (int e = 0; e < es; e++) hmf[(size_t)t * es + e] = fock[e];
      s.solve(t, view(S), view(G), nthreads);
    }
    err = D->dyson_start(view(G), view(S), hmf, mu, beta, dt);
    if (err_out) *err_out = err;
    if (err < tol) return iter + 1;
  }
  return -1;
}

// DecompSimulation::do_tstp (src/molNEsim_decomp_impl.h:86-104): predictor,
// then corrector passes of {rho, HF, Sigma, Dyson}.  corr_tol = 0 reproduces the
// reference's fixed CorrSteps; corr_tol > 0 stops once the Dyson update norm
// drops below it.  Returns the number of corrector passes.
int orc_do_tstp(void *d, int naux, const double *Vija, int tstp, const orc_green *G, const orc_green *S, cplx *hmf,
                const double *h0, double mu, double beta, double dt, int corr_max, double corr_tol, double *err_out,
                int nthreads, int extrapolate) {
  auto *D = (orc::Dyson *)d;
  const int nao = D->nao, es = nao * nao;
  orc::GF2Decomp s(nao, naux, Vija);
  orc::Green g = view(G), sg = view(S);
  if (extrapolate) D->Extrapolate(tstp, g);
  std::vector<cplx> rho(es);
  int it = 0;
  for (; it < corr_max; it++) {
    // get_dm: rho = i sig G^<(t,t) (src/greens.h:56-65)
    const cplx *gl = g.lesptr(tstp, tstp);
    for (int e = 0; e < es; e++) rho[e] = cplx(0.0, 1.0 * g.sig) * gl[e];
    for (int e = 0; e < es; e++) hmf[(size_t)tstp * es + e] = h0[e];
    s.solve_HF(tstp, hmf, rho.data());
    s.solve(tstp, sg, g, nthreads);
    double err = D->dyson_step(tstp, g, sg, hmf, mu, beta, dt);
    if (err_out) *err_out = err;
    if (corr_tol > 0 && err < corr_tol) { it++; break; }
  }
  return it;
}

}  // extern "C"
