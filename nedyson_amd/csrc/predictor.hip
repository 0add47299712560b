// The timestep driver DecompSimulation::do_tstp (src/molNEsim_decomp_impl.h:86-104)
// with a choice of initial guess: the reference's polynomial predictor
// dyson::Extrapolate (src/dyson.cpp:47-117), or that prediction corrected by
// the learned network (mlp.hip) on the time points where the polynomial one
// is weakest.
//
// Where the correction acts.  Extrapolate continues G^R(T, T-d) and
// G^<(T-d, T) along the time diagonal.  Its error is largest (i) on the 16
// columns just after the quench, t = k+1..k+16, whose diagonals cross the
// kink that the change of Hamiltonian at t = k+1 puts into G, and (ii) next
// to the diagonal, t = T-15..T.  These 32 points of the retarded row and of
// the lesser column form the per-orbital-pair basis of the network: for
// every orbital pair (i,j) its input is the extrapolation error (converged
// minus extrapolated value, real and imaginary parts) at the 2 x 32 points
// of each of the last six steps (768 numbers), and its output is the error
// expected at step T (128 numbers), which is added to the extrapolated slice.
// Inputs and outputs are scaled per component and zone by the RMS of that
// group over the six steps (the quench zone's errors exceed the diagonal
// zone's by orders of magnitude).  The plain corrector
// iteration then runs unchanged and decides convergence; a corrector that
// fails (non-finite update, or no decrease) is rerun from the polynomial
// guess.
#include "predictor.h"
#include "dyson.h"
#include <cmath>
#include <algorithm>

namespace ned {
namespace {

constexpr int NZ = 16;           // points per zone
constexpr int NPT = 2 * NZ;      // points per component
constexpr int NF = 2 * NPT * 2;  // features per pair and step: (comp, point, re/im)

__device__ __forceinline__ int pred_t(int pt, int T, int k) { return pt < NZ ? k + 1 + pt : T - (NPT - 1) + pt; }

// Rrow, Lrow: time row T of the G^R and G^< storage; comp 0: G^R(T, t)
// (block t of Rrow), comp 1: G^<(t, T) (block t of Lrow)
__device__ __forceinline__ z2 *pred_block(z2 *Rrow, z2 *Lrow, int comp, int t, size_t es) {
  return (comp == 0 ? Rrow : Lrow) + (size_t)t * es;
}

// E[(comp, pt)] = current slice values at the predictor points
__global__ void kern_pred_snapshot(z2 *Rrow, z2 *Lrow, int T, int k, int n, z2 *E) {
  const size_t es = (size_t)n * n;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)2 * NPT * es) return;
  const int cp = (int)(i / es), e = (int)(i % es), comp = cp / NPT, pt = cp % NPT;
  E[i] = pred_block(Rrow, Lrow, comp, pred_t(pt, T, k), es)[e];
}

// R[e][f] = (G - E) at the points, f = (comp * NPT + pt) * 2 + re/im
__global__ void kern_pred_resid(z2 *Rrow, z2 *Lrow, int T, int k, int n, const z2 *E, float *R) {
  const size_t es = (size_t)n * n;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)2 * NPT * es) return;
  const int cp = (int)(i / es), e = (int)(i % es), comp = cp / NPT, pt = cp % NPT;
  const z2 g = pred_block(Rrow, Lrow, comp, pred_t(pt, T, k), es)[e];
  const z2 d = zsub(g, E[i]);
  R[(size_t)e * NF + (size_t)cp * 2] = (float)d.x;
  R[(size_t)e * NF + (size_t)cp * 2 + 1] = (float)d.y;
}

struct Slots {
  const float *r[6];  // residuals of steps T-1 .. T-6
};

// X[e] = [R_{T-1}[e], ..., R_{T-6}[e]], each of the four groups of 32
// features (component x zone, f / 32) divided by its RMS over the six steps:
// the two zones differ by orders of magnitude and each keeps its own scale
constexpr int NG = 4, GF = NF / NG;
__global__ void kern_pred_features(Slots sl, int n2, float *X, float *sig) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n2) return;
  float sg[NG];
#pragma unroll
  for (int g = 0; g < NG; g++) {
    double s = 0;
#pragma unroll
    for (int q = 0; q < 6; q++)
      for (int f = g * GF; f < (g + 1) * GF; f++) {
        const double v = sl.r[q][(size_t)e * NF + f];
        s += v * v;
      }
    sg[g] = (float)std::sqrt(s / (6.0 * GF)) + 1e-30f;
    sig[(size_t)e * NG + g] = sg[g];
  }
#pragma unroll
  for (int q = 0; q < 6; q++)
    for (int f = 0; f < NF; f++) X[(size_t)e * 6 * NF + q * NF + f] = sl.r[q][(size_t)e * NF + f] / sg[f / GF];
}

// G += sigma * Y at the points
__global__ void kern_pred_apply(z2 *Rrow, z2 *Lrow, int T, int k, int n, const float *Y, const float *sig) {
  const size_t es = (size_t)n * n;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)2 * NPT * es) return;
  const int cp = (int)(i / es), e = (int)(i % es), comp = cp / NPT, pt = cp % NPT;
  z2 *g = pred_block(Rrow, Lrow, comp, pred_t(pt, T, k), es) + e;
  const int f = cp * 2;
  const float s = sig[(size_t)e * NG + f / GF];
  g->x += (double)(s * Y[(size_t)e * NF + f]);
  g->y += (double)(s * Y[(size_t)e * NF + f + 1]);
}

inline unsigned nb(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

int predictor_features() { return 6 * NF; }
int predictor_outputs() { return NF; }

void set_predictor(Ctx &c, Mlp *m) {
  if (m) {
    require(m->n_in == 6 * NF && m->n_out == NF, "predictor network must map 768 inputs to 128 outputs");
    require(m->device == c.device, "predictor network lives on another device");
  }
  c.pred_mlp = m;
}

const float *predictor_residual(Ctx &c, int tstp) {
  for (int s = 0; s < 7; s++)
    if (c.pred_ring_T[s] == tstp) return c.pred_ring.as<float>() + (size_t)s * c.es * NF;
  return nullptr;
}

void do_tstp(Ctx &c, int T, const double *h0, double mu, double beta, double dt, Prec prec, int corr_max,
             double corr_tol, int *iters_out, double *err_out) {
  const size_t es = c.es;
  const bool track = T >= c.k + 1 + NPT;  // the two zones are disjoint
  if (track) {
    c.pred_E.ensure(2 * NPT * es * sizeof(z2));
    c.pred_ring.ensure(7 * es * NF * sizeof(float));
  }
  // graph mode: the first pass runs as launched (it allocates scratch, tunes
  // tiles and fills the caches of the step); the second is captured on the
  // context's stream -- with the Dyson step's side streams forked from and
  // joined back into it -- and every later pass replays that graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  struct GraphGuard {
    hipGraph_t &g;
    hipGraphExec_t &x;
    ~GraphGuard() {
      if (x) (void)hipGraphExecDestroy(x);
      if (g) (void)hipGraphDestroy(g);
    }
  } gguard{graph, exec};
  auto graphed_pass = [&](double *e) {
    if (!exec) {
      hf_upload_h0(c, h0);
      NED_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
      c.capturing = true;
      try {
        hf_kernels(c, T);
        sigma_gf2(c, T, prec);
        dyson_step(c, T, mu, beta, dt, nullptr);
      } catch (...) {
        c.capturing = false;
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(c.stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      c.capturing = false;
      NED_HIP(hipStreamEndCapture(c.stream, &graph));
      NED_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    touch_time(c, T);
    NED_HIP(hipGraphLaunch(exec, c.stream));
    dyson_read_err(c, e);
  };
  auto corrector = [&](double &err, int &it, double &first) -> bool {
    double e[3] = {0, 0, 0};
    it = 0;
    err = 0;
    first = -1;
    for (; it < corr_max;) {
      if (c.graph_mode && it >= 1) {
        graphed_pass(e);
      } else {
        hf_decomp(c, T, h0);
        sigma_gf2(c, T, prec);
        touch_time(c, T);
        dyson_step(c, T, mu, beta, dt, e);
      }
      err = e[0] + e[1] + e[2];
      if (first < 0) first = err;
      it++;
      if (!std::isfinite(err)) return false;
      if (corr_tol > 0 && err < corr_tol) break;
    }
    return true;
  };
  extrapolate(c, T);
  z2 *Rrow = c.G.ret.as<z2>() + g_row(c, T) * es, *Lrow = c.G.les.as<z2>() + g_row(c, T) * es;
  if (track) {
    hipLaunchKernelGGL(kern_pred_snapshot, dim3(nb((long long)2 * NPT * es)), dim3(256), 0, c.stream, Rrow, Lrow, T, c.k, c.nao, c.pred_E.as<z2>());
    check_launch("predictor snapshot");
  }
  bool use_mlp = track && c.pred_mlp != nullptr;
  Slots sl{};
  for (int q = 0; use_mlp && q < 6; q++) {
    sl.r[q] = predictor_residual(c, T - 1 - q);
    if (!sl.r[q]) use_mlp = false;
  }
  if (use_mlp) {
    Mlp &m = *c.pred_mlp;
    c.pred_X.ensure(es * 6 * NF * sizeof(float));
    c.pred_Y.ensure(es * NF * sizeof(float));
    c.pred_sig.ensure(es * NG * sizeof(float));
    hipLaunchKernelGGL(kern_pred_features, dim3(nb((long long)es)), dim3(256), 0, c.stream, sl, (int)es,
                       c.pred_X.as<float>(), c.pred_sig.as<float>());
    m.infer(c.pred_X.as<float>(), (int)es, c.pred_Y.as<float>(), c.stream);
    hipLaunchKernelGGL(kern_pred_apply, dim3(nb((long long)2 * NPT * es)), dim3(256), 0, c.stream, Rrow, Lrow, T, c.k, c.nao, c.pred_Y.as<float>(), c.pred_sig.as<float>());
    check_launch("predictor apply");
    c.pred_used++;
  }
  double err = 0, first = 0;
  int it = 0;
  bool ok = corrector(err, it, first);
  if (use_mlp && (!ok || (corr_tol > 0 && err >= corr_tol && err >= first))) {
    // the learned guess did not lead to convergence: start again from the
    // polynomial one
    c.pred_fallback++;
    extrapolate(c, T);
    int it2 = 0;
    ok = corrector(err, it2, first);
    it += it2;
  }
  if (!ok) throw std::runtime_error("non-finite update in corrector");
  if (track) {
    const int slot = T % 7;
    hipLaunchKernelGGL(kern_pred_resid, dim3(nb((long long)2 * NPT * es)), dim3(256), 0, c.stream, Rrow, Lrow, T, c.k, c.nao, c.pred_E.as<z2>(),
                       c.pred_ring.as<float>() + (size_t)slot * es * NF);
    check_launch("predictor residual");
    c.pred_ring_T[slot] = T;
  }
  if (iters_out) *iters_out = it;
  if (err_out) *err_out = err;
}

void batch_do_tstp(DysonGroup &g, const std::vector<Ctx *> &cs, int T, const double *const *h0, const double *mu,
                   const double *beta, const double *dt, Prec prec, int corr_max, double corr_tol, int *iters,
                   double *err, int *status) {
  const int S = (int)cs.size();
  std::vector<int> live, its(S, 0);
  std::vector<double> errs(S, 0.0);
  for (int i = 0; i < S; i++) {
    Ctx &c = *cs[i];
    if (status) status[i] = 0;
    if (!(T > c.k && T <= c.nt)) {  // this system cannot take the step
      if (status) status[i] = 1;    // NED_ERR_ARG
      continue;
    }
    extrapolate(c, T);
    hf_upload_h0(c, h0[i]);
    live.push_back(i);
  }
  for (int pass = 0; pass < corr_max && !live.empty(); pass++) {
    std::vector<Ctx *> group;
    std::vector<double> m, b, d, e(3 * live.size());
    for (int i : live) {
      Ctx &c = *cs[i];
      hf_kernels(c, T);
      sigma_gf2(c, T, prec);
      touch_time(c, T);
      group.push_back(&c);
      m.push_back(mu[i]);
      b.push_back(beta[i]);
      d.push_back(dt[i]);
    }
    dyson_step_group(g, group, T, m.data(), b.data(), d.data(), e.data());
    std::vector<int> next;
    for (size_t q = 0; q < live.size(); q++) {
      const int i = live[q];
      errs[i] = e[3 * q] + e[3 * q + 1] + e[3 * q + 2];
      its[i]++;
      if (!std::isfinite(errs[i])) {
        if (status) status[i] = 4;  // NED_ERR_NONFINITE
        continue;
      }
      if (corr_tol > 0 && errs[i] < corr_tol) continue;
      next.push_back(i);
    }
    live.swap(next);
  }
  for (Ctx *c : cs) NED_HIP(hipStreamSynchronize(c->stream));
  for (int i = 0; i < S; i++) {
    if (iters) iters[i] = its[i];
    if (err) err[i] = errs[i];
  }
}

}  // namespace ned
