// C++ host mirror of the reference's GF2 timestep interface over the C ABI of
// nedyson.h.  The classes keep the reference's names, argument order and
// meaning so that a caller written against NEdyson reads the same:
//
//   NEdyson                                   nedyson_amd
//   GREEN G(nt, ntau, nao, -1)                 System sys(...); sys.G()
//   molGF2SolverDecomp(Vija, Viaj)             molGF2SolverDecomp(sys, Vija, Viaj)
//   solver.solve(tstp, Sigma, G)               solver.solve(tstp, sys.Sigma(), sys.G())
//   solver.solve_HF(tstp, hmf, rho)            solver.solve_HF(tstp, sys.hmf(), h0)
//   dyson(nt, ntau, nao, k, mode)              dyson(sys)
//   Dyson.dyson_step(n, G, Sig, hmf, mu, ...)  Dyson.dyson_step(n, G, Sig, hmf, mu, ...)
//   Dyson.Extrapolate(n, G)                    Dyson.Extrapolate(n, G)
//   Dyson.dyson_start(G, Sig, hmf, mu, ...)    Dyson.dyson_start(G, Sig, hmf, mu, ...)
//   DecompSimulation::do_boot / do_tstp        DecompSimulation::do_boot / do_tstp
//
// Differences, all forced by the storage living in HBM:
//   * GREEN here is a handle to one contour function of a System (the device
//     context); its blocks are copied in and out with get_*/set_* and
//     read/write, in the packing of green_func (src/greens.cpp:177-195).
//   * The solvers hold their System by reference (the reference's solvers
//     hold V by reference, src/mol2bSolver.h:80-88); every GREEN passed to
//     them must belong to that System.
//   * Errors throw nedyson_amd::Error carrying the ned_status instead of
//     assert()/exit().
#ifndef NEDYSON_HPP
#define NEDYSON_HPP

#include <complex>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "nedyson.h"

namespace nedyson_amd {

using cplx = std::complex<double>;

class Error : public std::runtime_error {
 public:
  Error(ned_status s, const std::string &msg) : std::runtime_error(msg), status(s) {}
  ned_status status;
};

inline void check(ned_status s) {
  if (s != NED_OK) throw Error(s, std::string("nedyson_amd: ") + ned_last_error());
}

inline int prec_of(const std::string &name) {
  if (name == "f64") return NED_PREC_F64;
  if (name == "f32") return NED_PREC_F32;
  if (name == "bf16x3") return NED_PREC_BF16X3;
  throw Error(NED_ERR_ARG, "nedyson_amd: unknown precision " + name);
}

class System;

// One contour function (G or Sigma) of a System.  Block addressing follows
// green_func: retptr(i,j) (j <= i), lesptr(i,j) (i <= j), tvptr(i,m), matptr(m).
class GREEN {
 public:
  int nt() const { return nt_; }
  int ntau() const { return ntau_; }
  int nao() const { return nao_; }
  int sig() const { return sig_; }
  size_t es() const { return (size_t)nao_ * nao_; }
  const System *owner() const { return sys_; }

  static long long retidx(int i, int j) { return (long long)i * (i + 1) / 2 + j; }
  static long long lesidx(int i, int j) { return (long long)j * (j + 1) / 2 + i; }
  long long tvidx(int i, int m) const { return (long long)i * (ntau_ + 1) + m; }

  // host copies of nblocks consecutive packed blocks of one component
  void read(ned_comp comp, long long first, long long nblocks, cplx *host) const {
    check(ned_green_read(ctx_, which_, comp, first, nblocks, host));
  }
  void write(ned_comp comp, long long first, long long nblocks, const cplx *host) {
    check(ned_green_write(ctx_, which_, comp, first, nblocks, host));
  }
  void get_ret(int i, int j, cplx *M) const { read(NED_RET, retidx(i, j), 1, M); }
  void get_les(int i, int j, cplx *M) const { read(NED_LES, lesidx(i, j), 1, M); }
  void get_tv(int i, int m, cplx *M) const { read(NED_TV, tvidx(i, m), 1, M); }
  void get_mat(int m, cplx *M) const { read(NED_MAT, m, 1, M); }
  void set_ret(int i, int j, const cplx *M) { write(NED_RET, retidx(i, j), 1, M); }
  void set_les(int i, int j, const cplx *M) { write(NED_LES, lesidx(i, j), 1, M); }
  void set_tv(int i, int m, const cplx *M) { write(NED_TV, tvidx(i, m), 1, M); }
  void set_mat(int m, const cplx *M) { write(NED_MAT, m, 1, M); }

  // green_func::get_dm (src/greens.h:56-65): rho = i sig G^<(t,t)
  void get_dm(int tstp, cplx *rho) const {
    get_les(tstp, tstp, rho);
    for (size_t e = 0; e < es(); e++) rho[e] *= cplx(0.0, (double)sig_);
  }

 private:
  friend class System;
  GREEN(const System *sys, ned_ctx *ctx, int which, int nt, int ntau, int nao, int sig)
      : sys_(sys), ctx_(ctx), which_(which), nt_(nt), ntau_(ntau), nao_(nao), sig_(sig) {}
  const System *sys_;
  ned_ctx *ctx_;
  int which_;
  int nt_, ntau_, nao_, sig_;
  friend class molGF2SolverDecomp;
  friend class molGF2Solver;
  friend class dyson;
};

// The mean field hmf(t), t = 0..nt (the reference's ZTensor<3> hmf).
class Hmf {
 public:
  void read(int first, int count, cplx *host) const { check(ned_hmf_read(ctx_, first, count, host)); }
  void write(int first, int count, const cplx *host) { check(ned_hmf_write(ctx_, first, count, host)); }
  const System *owner() const { return sys_; }

 private:
  friend class System;
  Hmf(const System *sys, ned_ctx *ctx) : sys_(sys), ctx_(ctx) {}
  const System *sys_;
  ned_ctx *ctx_;
};

// All device storage of one propagated system: G, Sigma, hmf, the
// interaction and scratch (the GREEN/ZTensor members of DecompSimulation,
// src/molNEsim_decomp_impl.h:19-39), on one GPU.
class System {
 public:
  System(int nt, int ntau, int nao, int naux, int k, int device = 0, int sig = -1) {
    d_.nao = nao;
    d_.naux = naux;
    d_.nt = nt;
    d_.ntau = ntau;
    d_.k = k;
    d_.sig = sig;
    check(ned_ctx_create(&ctx_, &d_, device));
  }
  ~System() {
    if (ctx_) ned_ctx_destroy(ctx_);
  }
  System(const System &) = delete;
  System &operator=(const System &) = delete;

  ned_ctx *ctx() const { return ctx_; }
  const ned_dims &dims() const { return d_; }
  GREEN G() { return GREEN(this, ctx_, NED_G, d_.nt, d_.ntau, d_.nao, d_.sig); }
  GREEN Sigma() { return GREEN(this, ctx_, NED_SIGMA, d_.nt, d_.ntau, d_.nao, d_.sig); }
  Hmf hmf() { return Hmf(this, ctx_); }
  void sync() { check(ned_sync(ctx_)); }

 private:
  ned_dims d_{};
  ned_ctx *ctx_ = nullptr;
};

namespace detail {
inline void same(const System *a, const System *b, const char *what) {
  if (a != b) throw Error(NED_ERR_ARG, std::string("nedyson_amd: ") + what + " belongs to another System");
}
}  // namespace detail

// molGF2SolverDecomp (src/mol2bSolver.h:78-123).  Vija is (nao, nao, naux)
// and Viaj (nao, naux, nao), row-major, as the reference's DTensor<3>; Viaj
// holds the same numbers and is checked against Vija when given.
class molGF2SolverDecomp {
 public:
  molGF2SolverDecomp(System &sys, const double *Vija, const double *Viaj = nullptr) : sys_(&sys) {
    const int n = sys.dims().nao, a = sys.dims().naux;
    std::vector<double> B((size_t)a * n * n);
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++)
        for (int Q = 0; Q < a; Q++) {
          const double v = Vija[((size_t)i * n + j) * a + Q];
          if (Viaj && Viaj[((size_t)i * a + Q) * n + j] != v)
            throw Error(NED_ERR_ARG, "nedyson_amd: Viaj(i,Q,j) differs from Vija(i,j,Q)");
          B[((size_t)Q * n + i) * n + j] = v;
        }
    check(ned_set_coulomb(sys.ctx(), B.data()));
  }

  // solve(tstp, Sigma, G) (src/mol2bSolver_decomp_impl.h:412-429); prec picks
  // the arithmetic of the contractions (Sigma is stored in FP64 always)
  void solve(int tstp, GREEN &Sigma, const GREEN &G, int prec = NED_PREC_F64) const {
    detail::same(Sigma.owner(), sys_, "Sigma");
    detail::same(G.owner(), sys_, "G");
    if (Sigma.which_ != NED_SIGMA || G.which_ != NED_G)
      throw Error(NED_ERR_ARG, "nedyson_amd: solve(tstp, Sigma, G) takes the System's Sigma and G");
    check(ned_sigma_gf2_decomp(sys_->ctx(), tstp, prec));
  }

  // hmf[tstp] = h0; rho = G.get_dm(tstp); solve_HF(tstp, hmf, rho)
  // (src/molNEsim_decomp_impl.h:92-94, src/molHFSolver.cpp:81-110)
  void solve_HF(int tstp, Hmf &hmf, const double *h0) const {
    detail::same(hmf.owner(), sys_, "hmf");
    check(ned_hf_decomp(sys_->ctx(), tstp, h0));
  }

 private:
  System *sys_;
};

// molGF2Solver (src/mol2bSolver.h:178-209): the full, non-decomposed second-Born
// self-energy, contracted through U_ijkl = sum_Q V_ijQ V_klQ (built on the
// device from the System's Cholesky factor, which must be set: construct a
// molGF2SolverDecomp first, or call ned_set_coulomb).
class molGF2Solver {
 public:
  explicit molGF2Solver(System &sys) : sys_(&sys) {}
  // solve(tstp, Sigma, G) (src/mol2bSolver_impl.h:5-120, 122)
  void solve(int tstp, GREEN &Sigma, const GREEN &G, int prec = NED_PREC_F64) const {
    detail::same(Sigma.owner(), sys_, "Sigma");
    detail::same(G.owner(), sys_, "G");
    if (Sigma.which_ != NED_SIGMA || G.which_ != NED_G)
      throw Error(NED_ERR_ARG, "nedyson_amd: solve(tstp, Sigma, G) takes the System's Sigma and G");
    check(ned_sigma_gf2_full(sys_->ctx(), tstp, prec));
  }

 private:
  System *sys_;
};

// dyson (src/dyson.h:16-140) for the System's dimensions.  set_matsubara
// must be called once G^M is written (it builds the mixing kernel that the
// reference's DecompSimulation builds with set_conv_tensor).
class dyson {
 public:
  explicit dyson(System &sys) : sys_(&sys) {}

  void set_matsubara(double beta) { check(ned_set_matsubara(sys_->ctx(), beta)); }

  // dyson_step(n, G, Sig, hmf, mu, beta, dt) (src/dyson_step.h:329-382);
  // err (3 doubles, optional) receives the ret, tv and les update norms
  void dyson_step(int n, GREEN &G, const GREEN &Sig, const Hmf &hmf, double mu, double beta, double dt,
                  double *err = nullptr) const {
    detail::same(G.owner(), sys_, "G");
    detail::same(Sig.owner(), sys_, "Sig");
    detail::same(hmf.owner(), sys_, "hmf");
    check(ned_dyson_step(sys_->ctx(), n, mu, beta, dt, err));
  }

  // dyson_start(G, Sig, hmf, mu, beta, dt) (src/dyson_boot.h:195-221)
  double dyson_start(GREEN &G, const GREEN &Sig, const Hmf &hmf, double mu, double beta, double dt) const {
    detail::same(G.owner(), sys_, "G");
    detail::same(Sig.owner(), sys_, "Sig");
    detail::same(hmf.owner(), sys_, "hmf");
    double err = 0;
    check(ned_dyson_start(sys_->ctx(), mu, beta, dt, &err));
    return err;
  }

  // Extrapolate(n, G) (src/dyson.cpp:47-117)
  void Extrapolate(int n, GREEN &G) const {
    detail::same(G.owner(), sys_, "G");
    check(ned_extrapolate(sys_->ctx(), n));
  }

  // G0_from_h0(G, mu, H0, beta, h) (src/dyson_free.h:226-297), steps 0..k
  void G0_from_h0(GREEN &G, double mu, const double *H0, double beta, double h) const {
    detail::same(G.owner(), sys_, "G");
    check(ned_g0_from_h0(sys_->ctx(), sys_->dims().k, H0, mu, beta, h));
  }

  // energy_conv(tstp, Sig, G, beta, dt) (src/dyson_integrals.h:24-70)
  double energy_conv(int tstp, const GREEN &Sig, const GREEN &G, double beta, double dt) const {
    detail::same(G.owner(), sys_, "G");
    detail::same(Sig.owner(), sys_, "Sig");
    std::vector<double> zero((size_t)sys_->dims().nao * sys_->dims().nao, 0.0);
    double out[3];
    check(ned_energy(sys_->ctx(), tstp, zero.data(), beta, dt, out));
    return 0.5 * out[1];
  }

 private:
  System *sys_;
};

// The equilibrium solve the reference delegates to gfmol::DecompSimulation
// (constructed at src/molNEsim_decomp_impl.h:21, run by do_mat at :50-52,
// read through mu() and fock() at :44-46 and sampled on the collocation
// nodes at :145-165): run() leaves G^M and Sigma^M in the System and keeps mu
// and the Fock matrix for DecompSimulation.  mode NED_MAT_GF2 or NED_MAT_HF;
// nelec > 0 solves for mu, otherwise mu stays at the constructor's value.
class MatsubaraSolver {
 public:
  MatsubaraSolver(System &sys, std::vector<double> h1, double nelec, double beta, int mode = NED_MAT_GF2,
                  double mu = 0.0)
      : sys_(&sys), h1_(std::move(h1)), fock_(h1_.size()), nelec_(nelec), beta_(beta), mu_(mu), mode_(mode) {}

  // do_mat: returns the outer iterations used; throws on NED_ERR_NOCONV
  int run(int maxiter, double tol) {
    int it = 0;
    check(ned_matsubara_gf2(sys_->ctx(), h1_.data(), nelec_ > 0 ? nelec_ : 0.0, beta_, mode_, nelec_ > 0 ? 0 : 1, &mu_,
                            maxiter, tol, fock_.data(), &it, &err_));
    return it;
  }
  double mu() const { return mu_; }
  const std::vector<double> &fock() const { return fock_; }
  double beta() const { return beta_; }
  double last_update() const { return err_; }

 private:
  System *sys_;
  std::vector<double> h1_, fock_;
  double nelec_, beta_, mu_, err_ = 0;
  int mode_;
};

// DecompSimulation's propagation loop (src/molNEsim_decomp_impl.h:63-144)
// on a System.  h0 is the post-quench one-body Hamiltonian, fock the
// equilibrium mean field; both real nao x nao.
class DecompSimulation {
 public:
  DecompSimulation(System &sys, std::vector<double> h0, std::vector<double> fock, double mu, double beta, double dt,
                   int CorrSteps, int BootMax, double BootTol, int prec = NED_PREC_F64)
      : sys_(&sys), h0_(std::move(h0)), fock_(std::move(fock)), mu_(mu), beta_(beta), dt_(dt),
        CorrSteps_(CorrSteps), BootMax_(BootMax), BootTol_(BootTol), prec_(prec) {}

  // do_boot (src/molNEsim_decomp_impl.h:63-82): passes iter = 0..BootMax
  // until the update norm is below BootTol; bootstrap_converged records
  // whether it was reached, as in the reference.  Returns the passes used.
  int do_boot() {
    int iters = 0;
    double err = 0;
    ned_status s = ned_do_boot(sys_->ctx(), fock_.data(), mu_, beta_, dt_, prec_, BootMax_, BootTol_, &iters, &err);
    bootstrap_converged = s == NED_OK;
    if (s != NED_ERR_NOCONV) check(s);
    return iters;
  }

  // do_tstp (src/molNEsim_decomp_impl.h:86-104): predictor, then CorrSteps
  // passes of {rho, HF, Sigma, Dyson}
  void do_tstp(int tstp) {
    check(ned_do_tstp(sys_->ctx(), tstp, h0_.data(), mu_, beta_, dt_, prec_, CorrSteps_, 0.0, nullptr, nullptr));
  }

  // do_energy (src/molNEsim_decomp_impl.h:132-144) at one time step:
  // {eKin, ePot, particle number}
  void energy(int tstp, double *out3) const {
    check(ned_energy(sys_->ctx(), tstp, h0_.data(), beta_, dt_, out3));
  }

  bool bootstrap_converged = false;

 private:
  System *sys_;
  std::vector<double> h0_, fock_;
  double mu_, beta_, dt_;
  int CorrSteps_, BootMax_;
  double BootTol_;
  int prec_;
};

}  // namespace nedyson_amd

#endif
