/*
 * nedyson_amd C ABI: the Kadanoff-Baym GF2 timestep of NEdyson on one MI355X
 * (gfx950) GPU.
 *
 * One context owns one propagated system: the contour functions G and Sigma
 * in the packing of the reference's green_func (src/greens.cpp:177-195), the
 * mean field hmf, the Cholesky-factored interaction and all scratch, resident
 * in HBM, plus its HIP streams (the context's stream, and two more on which
 * the Dyson step runs its bulk history updates beside the dependency chain
 * and its mixing row beside the retarded row; every call's work is joined
 * back into the context's stream before the call returns or the next phase
 * starts).  Calls on different contexts are independent;
 * calls on one context are serialised by the caller (the reference's solver
 * objects are likewise not re-entrant, src/mol2bSolver.h:80-88).
 *
 * Storage (blocks are row-major nao x nao complex128, es = nao*nao):
 *   ret  G^R(i, j<=i)   block i(i+1)/2 + j
 *   les  G^<(i<=j, j)   block j(j+1)/2 + i
 *   tv   G^rceil(i, m)  block i*(ntau+1) + m
 *   mat  G^M(m)         block m
 * Imaginary time runs over the ntau+1 Gauss-Lobatto-Legendre nodes
 * tau_m = beta (x_m + 1)/2 (the reference's cpsitop collocation,
 * src/dyson.cpp:40).
 *
 * Every entry point returns NED_OK or an error code; ned_last_error() gives the
 * message.  Nothing here falls back to the CPU.
 */
#ifndef NEDYSON_H
#define NEDYSON_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  NED_OK = 0,
  NED_ERR_ARG = 1,       /* bad sizes, indices or unsupported input */
  NED_ERR_NOMEM = 2,     /* device allocation failed */
  NED_ERR_DEVICE = 3,    /* HIP runtime / kernel error */
  NED_ERR_NONFINITE = 4, /* a NaN or Inf appeared in the solution */
  NED_ERR_NOCONV = 5,    /* iteration did not reach its tolerance */
  NED_ERR_STATE = 6      /* call out of order (e.g. no interaction set) */
} ned_status;

typedef enum { NED_PREC_F64 = 0, NED_PREC_F32 = 1, NED_PREC_BF16X3 = 2 } ned_prec;

typedef enum { NED_G = 0, NED_SIGMA = 1 } ned_which;
typedef enum { NED_RET = 0, NED_LES = 1, NED_TV = 2, NED_MAT = 3 } ned_comp;
typedef enum { NED_MAT_HF = 0, NED_MAT_GF2 = 1 } ned_mat_mode;

typedef struct {
  int nao;   /* orbitals */
  int naux;  /* Cholesky vectors (nalpha) */
  int nt;    /* last time step stored */
  int ntau;  /* imaginary-time nodes are 0..ntau */
  int k;     /* Gregory / BDF order */
  int sig;   /* -1 fermions */
} ned_dims;

typedef struct ned_ctx ned_ctx;

const char *ned_last_error(void);

/* Replaces the GREEN(nt, ntau, nao, -1) allocations of DecompSimulation
 * (src/molNEsim_decomp_impl.h:33,39) and the dyson object
 * (src/dyson.cpp:25-43): all device storage of one system. */
ned_status ned_ctx_create(ned_ctx **out, const ned_dims *dims, int device);
ned_status ned_ctx_destroy(ned_ctx *ctx);
ned_status ned_sync(ned_ctx *ctx);
/* the sizes the context was created with (GREEN::nt(), ntau(), size1()) */
ned_status ned_ctx_get_dims(const ned_ctx *ctx, ned_dims *out);

/* Storage of the time histories, chosen at creation (no counterpart in the
 * reference, whose green_func holds every component whole in FP64,
 * src/greens.cpp:177-195).  For runs whose FP64 triangles exceed the GPU's
 * memory (nao = 64, nt = 2000: four 131 GB triangles):
 *  NED_STORE_SIGMA_F32: the Sigma^R / Sigma^< triangles hold complex FP32;
 *    the Dyson step and the energy widen every element to FP64 as they read
 *    it (Sigma^rceil, Sigma^M and all of G stay FP64).  The Sigma
 *    contractions round their result once more, to FP32.
 *  NED_STORE_G_RING: G^R / G^< keep only their last k + 3 time rows (a
 *    timestep reads rows T - k - 1 .. T, src/dyson.cpp:64-116 and
 *    src/dyson_step.h; Sigma at T reads row T alone).  ned_green_read of a
 *    row no longer held, and any call that needs one, fail with NED_ERR_ARG;
 *    ned_green_write of rows of G^R / G^< makes its last row the newest.
 *    ned_tti_extend of G is not available.
 * Results equal those of whole FP64 storage: bit for bit for the ring; for
 * FP32 Sigma within the FP32 rounding of Sigma. */
enum { NED_STORE_SIGMA_F32 = 1, NED_STORE_G_RING = 2 };
ned_status ned_ctx_create_ex(ned_ctx **out, const ned_dims *dims, int device, unsigned flags);
/* the storage flags, and the first and last time row of G^R / G^< held
 * (0 and nt for whole storage) */
ned_status ned_ctx_get_storage(const ned_ctx *ctx, unsigned *flags, int *g_first, int *g_last);

/* Replaces molGF2SolverDecomp(Vija, Viaj) (src/mol2bSolver.h:101-111):
 * B_Qij is naux x nao x nao, B[Q][i][j] = Vija(i,j,Q); must be symmetric in
 * (i,j).  At nao = 64 that symmetry also splits the exchange contraction into
 * its parts even and odd in the contracted pair (half the flops); the
 * environment variable NED_SYMX=0, read here, keeps the full contraction. */
ned_status ned_set_coulomb(ned_ctx *ctx, const double *B_Qij);

/* Copy nblocks packed blocks starting at first_block between host and device
 * (green_func::lesptr/retptr/tvptr/matptr addressing, src/greens.cpp:177-195). */
ned_status ned_green_write(ned_ctx *ctx, int which, int comp, long long first_block, long long nblocks,
                           const void *host_blocks);
ned_status ned_green_read(ned_ctx *ctx, int which, int comp, long long first_block, long long nblocks,
                          void *host_blocks);
/* mean field hmf(t) for t in [first, first+count), complex128 nao x nao each */
ned_status ned_hmf_write(ned_ctx *ctx, int first, int count, const void *host);
ned_status ned_hmf_read(ned_ctx *ctx, int first, int count, void *host);

/* molGF2SolverDecomp::solve(tstp, Sigma, G) (src/mol2bSolver_decomp_impl.h:412-429):
 * overwrites Sigma^R(T,0..T), Sigma^<(0..T,T) and Sigma^rceil(T,:) from the
 * same slices of G.  prec selects the arithmetic of the contractions; Sigma is
 * stored in FP64 in every case. */
ned_status ned_sigma_gf2_decomp(ned_ctx *ctx, int tstp, int prec);

/* molGF2Solver::solve(tstp, Sigma, G) (src/mol2bSolver_impl.h:5-120, 122;
 * class at src/mol2bSolver.h:178-209): the same slices as
 * ned_sigma_gf2_decomp from the full, non-decomposed interaction
 * U_ijkl = sum_Q B[Q][i][j] B[Q][k][l] and the exchange-combined
 * 2 U_{kj,qm} - U_{qj,km} (the reference's Uijkl_exch), both built on the
 * device from the context's Cholesky factor on first use (n^4 each):
 * 24 nao^5 flop per time point in four batched products. */
ned_status ned_sigma_gf2_full(ned_ctx *ctx, int tstp, int prec);

/* ned_sigma_gf2_decomp with no device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * context (G, Sigma, the interaction) is set up beforehand as usual; this call
 * then allocates nothing: its chunk scratch is a static device array of
 * ned_sigma_noalloc_scratch_bytes() (256 MiB) per GPU, held until the process
 * ends, and its split-bf16 GEMMs use the tile variant chosen earlier for the
 * shape (or a fixed default) without timing runs.  Time points are processed
 * in chunks that fit that array; a shape whose single time point needs more
 * returns NED_ERR_ARG.  Every call on one GPU shares that GPU's array, so calls
 * that overlap on one GPU (other threads, other contexts' streams) race:
 * serialise them.  Same results as ned_sigma_gf2_decomp up to the chunking
 * (bitwise in FP64). */
ned_status ned_sigma_gf2_decomp_noalloc(ned_ctx *ctx, int tstp, int prec);
size_t ned_sigma_noalloc_scratch_bytes(void);

/* G.get_dm(tstp, rho); hmf[tstp] = h0; molHFSolverDecomp::solve_HF(tstp, hmf, rho)
 * (src/greens.h:56-65, src/molHFSolver.cpp:81-110).  h0 is real nao x nao. */
ned_status ned_hf_decomp(ned_ctx *ctx, int tstp, const double *h0);

/* Equilibrium state: the gfmol Matsubara solve the reference runs before
 * propagating (gfmol::DecompSimulation, constructed at
 * src/molNEsim_decomp_impl.h:21 and run at :50-52; its G^M and Sigma^M are
 * sampled on the collocation nodes at :145-165, and its mu() and fock() feed
 * G0_from_h0 at :44-46).
 *   G^M = G0 + G0 o Sigma^M o G^M on the ntau+1 GLL nodes, with
 *   G0 = G0[F - mu], F = h1 + 2J[rho] - K[rho], rho = -G^M(beta), and
 *   Sigma^M the second-Born self-energy (mode NED_MAT_GF2) or 0 (NED_MAT_HF).
 * mu is found so that 2 tr rho = nelec, unless fix_mu is nonzero (then *mu_io
 * is used as given and nelec is ignored).  Starts from the G^M held in the
 * context, or from G0[h1] if that is zero.  Outputs: G^M and Sigma^M in the
 * context, the T = 0 slices of G and Sigma and hmf[0] consistent with them,
 * *mu_io, fock_out (nao x nao, the F of the solution), the number of outer
 * iterations and the last update max|dG^M|.  Returns NED_ERR_NOCONV if
 * max|dG^M| >= tol after maxiter iterations.  Real orbitals only (the
 * Coulomb factor is real symmetric); call ned_set_matsubara(beta) afterwards. */
ned_status ned_matsubara_gf2(ned_ctx *ctx, const double *h1, double nelec, double beta, int mode, int fix_mu,
                             double *mu_io, int maxiter, double tol, double *fock_out, int *iters_out,
                             double *err_out);

/* Imaginary-time setup: builds the mixing kernel of G^M (the decomposed
 * path's GMCT, src/greens.cpp:112-146, initialised here as the full path does,
 * src/molNEsim_impl.h:65) and the GLL mass matrix for beta.  Call after G^M is
 * written and whenever G^M or beta change. */
ned_status ned_set_matsubara(ned_ctx *ctx, double beta);

/* dyson::dyson_step(n, G, Sig, hmf, mu, beta, dt) (src/dyson_step.h:329-382).
 * err_out (3 doubles, may be NULL) receives the ret, tv and les update norms
 * (src/dyson_step.h:69,104,164,315). */
ned_status ned_dyson_step(ned_ctx *ctx, int tstp, double mu, double beta, double dt, double *err_out);

/* dyson::Extrapolate(n, G) (src/dyson.cpp:47-117). */
ned_status ned_extrapolate(ned_ctx *ctx, int tstp);

/* dyson::dyson_start(G, Sig, hmf, mu, beta, dt) (src/dyson_boot.h:195-221). */
ned_status ned_dyson_start(ned_ctx *ctx, double mu, double beta, double dt, double *err_out);

/* DecompSimulation::do_boot (src/molNEsim_decomp_impl.h:63-82): passes
 * iter = 0..maxiter (maxiter is the reference's BootMax) of
 * {hmf[0..k] = fock; Sigma(0..k); dyson_start}, stopping once err < tol.
 * *iters_out receives the passes used; NED_ERR_NOCONV if none reached tol
 * (the reference exits with status 0 there, src/molNEsim.cpp:129-131). */
ned_status ned_do_boot(ned_ctx *ctx, const double *fock, double mu, double beta, double dt, int prec, int maxiter,
                       double tol, int *iters_out, double *err_out);

/* DecompSimulation::do_tstp (src/molNEsim_decomp_impl.h:86-104): predictor,
 * then up to corr_max passes of {rho, HF, Sigma, Dyson}.  corr_tol = 0 runs
 * exactly corr_max passes (the reference's CorrSteps); corr_tol > 0 stops once
 * the Dyson update norm is below it.  *iters_out, *err_out may be NULL. */
ned_status ned_do_tstp(ned_ctx *ctx, int tstp, const double *h0, double mu, double beta, double dt, int prec,
                       int corr_max, double corr_tol, int *iters_out, double *err_out);

/* Graph mode of ned_do_tstp (no counterpart in the reference): with enable
 * nonzero, the corrector passes of a step after the first are one HIP graph,
 * captured from the second pass on the context's stream (the HF kernels, the
 * Sigma chain and the Dyson step with its side streams forked and joined) and
 * replayed for every later pass of that step; the first pass runs as
 * launched.  Same results as without it.  Also set by NED_GRAPH=1 in the
 * environment at ned_ctx_create. */
ned_status ned_set_graph(ned_ctx *ctx, int enable);

/* energy_conv(tstp, Sigma, G, beta, dt) (src/dyson_integrals.h:24-70) and the
 * kinetic part of do_energy (src/molNEsim_decomp_impl.h:132-144):
 * out[0] = Re tr(rho (hmf + h0)), out[1] = 2 * energy_conv, out[2] = 2 tr rho. */
ned_status ned_energy(ned_ctx *ctx, int tstp, const double *h0, double beta, double dt, double *out3);

/* dyson::G0_from_h0(G, mu, H0, beta, dt) (src/dyson_free.h:226-297) for
 * time steps 0..tmax (tmax = k is the reference's bootstrap initial guess):
 * G^R, G^<, G^rceil of the constant mean field fock (real nao x nao).  G^M is
 * not written, as in the reference. */
ned_status ned_g0_from_h0(ned_ctx *ctx, int tmax, const double *fock, double mu, double beta, double dt);

/* Time-translation-invariant continuation of slice tsrc to all t <= tsrc:
 * ret(t,t') = ret(tsrc, tsrc-(t-t')), les(t',t) = les(tsrc-(t-t'), tsrc),
 * tv(t,:) = tv(tsrc,:) -- the relation the bootstrap imposes on G^R
 * (src/dyson_boot.h:78-88).  Used to seed long equilibrium histories. */
ned_status ned_tti_extend(ned_ctx *ctx, int which, int tsrc);

/* Event timing of the hot phases on the context's stream.  ned_profile_read
 * fills 8 doubles: {ms, count} for the Sigma exchange GEMM (both launches of
 * the split form counted as one), the whole Sigma
 * and the whole Dyson step, then the algorithmic flops of the timed exchange
 * GEMM launches and a reserved 0, accumulated since the last read. */
ned_status ned_profile(ned_ctx *ctx, int enable);
ned_status ned_profile_read(ned_ctx *ctx, double *out8);

/* The tile variants of the split-bf16 NT GEMMs (the exchange and Pi
 * products), for tuning and tests; no counterpart in the reference.  The
 * first call for a new shape times every fitting variant and keeps the
 * fastest; NED_NT_VARIANT=<index> in the environment (read on every call)
 * pins one wherever it fits.  ned_gemm_variant_name returns a static string
 * ("?" out of range). */
int ned_gemm_variant_count(void);
const char *ned_gemm_variant_name(int v);

/* ---------------------------------------------------------------------------
 * Learned initial guess.  No counterpart in the reference: these replace the
 * polynomial predictor dyson::Extrapolate (src/dyson.cpp:47-117) inside
 * ned_do_tstp, on 2 x 32 time points per slice (the 16 columns after the
 * quench, t = k+1..k+16, and t = T-15..T, of the retarded row and the lesser
 * column), by adding the extrapolation error that a three-layer network
 * predicts from the errors of the last six steps.  The corrector iteration is
 * unchanged and decides convergence; a corrector that fails from the learned
 * guess is rerun from the polynomial one.
 *
 * Network: x (n_in) -> tanh(W1 x + b1) -> tanh(W2 . + b2) -> W3 . + b3 (n_out),
 * FP32, parameters flat in the order W1, b1, W2, b2, W3, b3 (weights row-major
 * out x in).  Training: minibatch Adam on the mean-squared error. */
typedef struct ned_mlp ned_mlp;
ned_status ned_mlp_create(ned_mlp **out, int n_in, int width, int n_out, int device, unsigned long long seed);
ned_status ned_mlp_destroy(ned_mlp *m);
ned_status ned_mlp_nparams(ned_mlp *m, long long *n);
ned_status ned_mlp_get_params(ned_mlp *m, float *out);
ned_status ned_mlp_set_params(ned_mlp *m, const float *in); /* also clears the Adam moments */
/* mean-squared loss of rows x n_in inputs X against rows x n_out targets T and
 * its gradient (nparams floats, may be NULL) */
ned_status ned_mlp_loss_grad(ned_mlp *m, const float *X, const float *T, int rows, double *loss, float *grads);
/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with the last gradient */
ned_status ned_mlp_adam_step(ned_mlp *m, float lr);
/* epochs of shuffled minibatches (seeded); loss_per_epoch may be NULL */
ned_status ned_mlp_train(ned_mlp *m, const float *X, const float *T, int rows, int epochs, int batch, float lr,
                         unsigned long long seed, float *loss_per_epoch);
ned_status ned_mlp_predict(ned_mlp *m, const float *X, int rows, float *Y);
/* inference (ned_mlp_predict and the timestep's guess) with the weights and
 * each layer's input rounded to bf16 and the products on the bf16 matrix
 * cores, FP32 accumulation, bias and tanh (on != 0), or all in FP32 (0, the
 * default); training is always FP32.  BF16 needs n_in % 64 == 0 and width,
 * n_out % 128 == 0.  No counterpart in the reference. */
ned_status ned_mlp_set_bf16(ned_mlp *m, int on);
/* initial guess of ned_do_tstp: NULL = polynomial only, else a network with
 * n_in = 768, n_out = 128 on the context's device.  The context keeps the
 * pointer: unset it (or destroy the context) before destroying the network,
 * and do not share one network between contexts that step concurrently. */
ned_status ned_set_predictor(ned_ctx *ctx, ned_mlp *m);
/* extrapolation error of step tstp at the predictor points, nao^2 x 128 floats
 * (pair (i,j) row-major, then (component, point, re/im)); the last seven
 * steps taken with T >= k + 33 are kept */
ned_status ned_predictor_residual(ned_ctx *ctx, int tstp, float *out);
/* steps that used the network, and of those, how many fell back */
ned_status ned_predictor_stats(ned_ctx *ctx, long long *used, long long *fallbacks);

/* ---------------------------------------------------------------------------
 * Independent systems stepped together.  No counterpart in the reference
 * (which propagates one system per process): a batch holds n existing
 * contexts, possibly of different nao, naux, nt and ntau, on the same or on
 * different devices, stepped at the same tstp.  Grouped (the default, all
 * contexts on one device): each corrector pass issues every system's HF and
 * Sigma on its own stream, then one Dyson step for all of them whose
 * retarded and lesser recurrences -- the dependency chains and bulk panel
 * updates that are most of its launches -- run as one launch per link and per
 * panel for the whole group (the system in the grid's third dimension); the
 * initial guess is the polynomial one, and with corr_tol > 0 a converged
 * system leaves the group while the others go on.  Otherwise (contexts on
 * several devices, or ned_batch_set_grouped(b, 0)) ned_do_tstp runs on each,
 * one host thread and stream per system.  The contexts stay owned by the
 * caller and must outlive the batch.  h0[i] is system i's real nao_i x nao_i
 * Hamiltonian; iters, err and status (arrays of n, each may be NULL) receive
 * the per-system results; the return value is the first failing system's
 * status. */
typedef struct ned_batch ned_batch;
ned_status ned_batch_create(ned_batch **out, ned_ctx *const *ctxs, int n);
ned_status ned_batch_destroy(ned_batch *b);
ned_status ned_batch_set_grouped(ned_batch *b, int grouped);
ned_status ned_batch_do_tstp(ned_batch *b, int tstp, const double *const *h0, const double *mu, const double *beta,
                             const double *dt, int prec, int corr_max, double corr_tol, int *iters, double *err,
                             int *status);

/* Gauss-Lobatto-Legendre nodes x_0..x_ntau on [-1,1] (ascending). Host-only. */
ned_status ned_gll_nodes(int ntau, double *x);

/* Integrator(k) tables (src/integration.h:23-124): poly_interp (k+1)^2,
 * poly_diff (k+1)^2, poly_integ (k+1)^3, bd_weights k+2, start (k+1)^2,
 * Omega (k+1)^2, ex_weights k+1.  Host-only. */
ned_status ned_integrator(int k, double *interp, double *diff, double *integ, double *bd, double *start,
                          double *Omega, double *ex);

#ifdef __cplusplus
}
#endif
#endif
