// Device context: one propagated system (G, Sigma, mean field, interaction)
// resident in HBM in the packing of the reference's green_func
// (src/greens.cpp:177-195): ret(i,j<=i) at block i(i+1)/2+j, les(i<=j,j) at
// block j(j+1)/2+i, tv(i,m) at block i*(ntau+1)+m, mat(m) at block m; every
// block a row-major nao x nao complex128 matrix.
#pragma once
#include "common.h"
#include <vector>
#include <memory>
#include <cstdio>
#include <stdexcept>

namespace ned {

struct Mlp;

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void ensure(size_t b) {
    if (b <= bytes) return;
    release();
    if (b == 0) return;
    NED_HIP(hipMalloc(&p, b));
    bytes = b;
  }
  template <typename T>
  T *as() const { return reinterpret_cast<T *>(p); }
};

struct Green {
  DevBuf ret, les, tv, mat;
};

// Shapes of the exchange contraction split by the symmetry of U' (nao = n,
// naux = a): ns pairs m <= q and na pairs m < q; the symmetric GEMM has K =
// ns + a (the Pi columns ride on it) padded to ks, the antisymmetric K = na
// padded to ka; operand A rows hold [sym | Pi | 0 | anti | 0] (stride kx); the
// outputs are padded to nsp / nap columns and share rows of stride ldh.
struct SymGeom {
  int ns, na, nsp, nap, ldh;
  long long ks, ka, kx;
};
inline SymGeom sym_geom(int n, int a) {
  SymGeom g;
  g.ns = n * (n + 1) / 2;
  g.na = n * (n - 1) / 2;
  g.ks = (g.ns + a + 63) / 64 * 64;
  g.ka = (g.na + 63) / 64 * 64;
  g.kx = g.ks + g.ka;
  g.nsp = (g.ns + 255) / 256 * 256;  // whole 256-column tiles
  g.nap = (g.na + 255) / 256 * 256;
  g.ldh = g.nsp + g.nap;
  return g;
}

struct Ctx {
  int nao = 0, naux = 0, nt = 0, ntau = 0, k = 0, sig = -1, device = 0;
  int es = 0, np = 0;
  hipStream_t stream = nullptr;
  // second stream of the Dyson recurrences (bulk panel updates overlapped
  // with the dependency chain), third stream of the Dyson step (the mixing
  // row, beside the retarded recurrence), and the events that order them
  hipStream_t side = nullptr, side2 = nullptr;
  hipEvent_t ev_main = nullptr, ev_side[2] = {nullptr, nullptr}, ev_fork = nullptr, ev_tv = nullptr;
  Green G, S;
  // storage of the time histories (fixed at creation):
  // s32: the Sigma^R / Sigma^< triangles hold complex FP32 (float2) elements
  //   (Sigma^rceil and Sigma^M stay FP64); every reader widens to FP64;
  // gring > 0: G^R / G^< keep only the last gring time rows, row t in slot
  //   t mod gring of nt + 1 blocks (the timestep reads rows T - k - 1 .. T);
  //   g_top is the newest row written, and rows g_top - gring + 1 .. g_top
  //   are the ones held
  bool s32 = false;
  int gring = 0, g_top = -1;
  DevBuf hmf;      // (nt+1) x es complex
  DevBuf Bq;       // naux x nao x nao double, B[Q][i][j] = V(i,j,Q)
  DevBuf Bq32;     // float copy
  DevBuf U64, U32; // (nao^2 x nao^2): U'[(m,q)][(k,j)] = sum_Q B[Q][k][m] B[Q][q][j]
  // operand B of the split-bf16 exchange GEMM, hi / lo planes of
  // Baug[(k,j)][c] = -U'[c][(k,j)] (c < nao^2), 2 B[c-nao^2][k][j] (c >= nao^2):
  // the bubble's Pi V product rides along the exchange as nao^2 + naux extra K
  DevBuf Uah, Ual;
  bool uaug = false;  // Baug built (nao^2 and 2 naux multiples of 128, naux of 64)
  bool fast64 = false;  // nao = 64 as well: the batched 64 x 64 kernels carry the chain
  DevBuf Vh, Vl;        // B[Q][i][l] split into bf16 hi / lo (fast64)
  // fast64 with the exchange split into its parts symmetric and antisymmetric
  // under (m,q) -> (q,m): the pair table (m << 8 | q) and the two right operands
  bool symx = false;
  DevBuf sym_pairs, Ush, Usl, Uxh, Uxl;
  // the same split on the FP64 / FP32 GEMM path (nao >= 8): its right operands
  // in FP64 and FP32 (both parts, row-major K x N blocks)
  bool symg = false;
  DevBuf Usg64, Usg32;
  // the full (non-decomposed) Sigma: U4[(i,l)][(n,p)] = U_{il,np} and
  // Wx[(k,q,m)][j] = 2 U_{kj,qm} - U_{qj,km} (built on first use)
  DevBuf U4_64, U4_32, Wx64, Wx32;
  bool have_coulomb = false;
  // Sigma scratch: the buffers of one chunk of time points, carved from one
  // allocation
  DevBuf sg_arena;
  int sigma_chunk = 0;  // items per chunk (0 = choose automatically)
  // Dyson recurrences: blocks per panel, and the largest number of later
  // blocks for which a panel's bulk update is split into 16-row/column bands
  int rpanel = 16, panel_band_max = 128;
  // the form of the recurrences' dependency chains (NED_CHAIN):
  //   1 (default): one launch per link, the panel's history sums carried
  //     times M^-1 so that a link is one product (kern_*_xform, kern_*_link),
  //     with transfer blocks formed once per time step (dy_tbr / dy_tbl, valid
  //     for step tb_T[0] / tb_T[1] at mu tb_mu, dt tb_dt and tb_rp blocks a panel);
  //   0: one launch per link of two products (kern_*_step_band), the form the
  //     grouped recurrences of many systems use;
  //   2: one launch per panel (kern_ret_chain / kern_les_chain), with hand-over
  //     flags (one per block and band, for each recurrence) and counters
  //     (ticket, done, epoch); at nao = 64, T = 1000 no faster than 0: its
  //     links wait on device-scope hand-overs about as long as a launch takes
  int chain_mode = 1;
  DevBuf dy_tbr, dy_tbl;
  int tb_T[2] = {-1, -1}, tb_rp[2] = {0, 0};
  double tb_mu[2] = {0, 0}, tb_dt[2] = {0, 0};
  DevBuf dy_flags, dy_ctr;
  // Dyson state
  DevBuf mixK;                // mixing kernel of G^M (np x np blocks)
  DevBuf mass;                // (np x np) GLL mass matrix
  DevBuf Minv;                // (nt+1) blocks: M_j^{-1} cache
  std::vector<char> minv_valid;
  double minv_mu = 0.0 / 0.0, minv_dt = 0.0 / 0.0;
  double mix_beta = 0;
  bool have_mix = false;
  DevBuf dy_tvA, dy_tvB, dy_tvpart, dy_tvinfo;  // the mixing row's own scratch
  DevBuf dy_info, dy_js, dy_gA, dy_gB, dy_err, dy_Q, dy_QA, dy_X, dy_M, dy_Qs, dy_part, dy_scr, dy_Y;
  DevBuf dy_la, dy_lb, dy_mixtmp, dy_h0, dy_xq, dy_bootI, dy_en, dy_hfpart;
  // inverses of the k x k block start-up systems: the retarded one depends on
  // times T-k..T-1 only (valid for every corrector pass of step retinv_T), the
  // lesser one on times 0..k only (valid until the early history changes)
  DevBuf dy_retinv, dy_lesinv, dy_Xs, dy_pinv, dy_d2;
  int retinv_T = -1;
  // initial-guess predictor: extrapolated values at the predictor points,
  // extrapolation errors of the last 7 steps (ring, slot T % 7 holds step
  // pred_ring_T[slot]), network inputs / outputs / scales
  Mlp *pred_mlp = nullptr;
  DevBuf pred_E, pred_ring, pred_X, pred_Y, pred_sig;
  int pred_ring_T[7] = {-1, -1, -1, -1, -1, -1, -1};
  long long pred_used = 0, pred_fallback = 0;
  DevBuf wtab_d;  // integrator tables of order wtab_k
  int wtab_k = -1;
  bool lesinv_valid = false;
  // event timing of the hot phases (enabled by ned_profile)
  bool prof = false;
  double prof_flops = 0;  // algorithmic flops of the timed exchange-GEMM launches
  std::vector<hipEvent_t> ev[6];  // U-GEMM start/end, Sigma start/end, Dyson start/end
  // corrector passes after the first replayed from one captured HIP graph per
  // step (ned_set_graph); capturing: a capture is open on the stream
  bool graph_mode = false, capturing = false;
  void mark(int slot) {
    if (!prof || capturing) return;
    hipEvent_t e;
    NED_HIP(hipEventCreate(&e));
    NED_HIP(hipEventRecord(e, stream));
    ev[slot].push_back(e);
  }
};

size_t tri(int t);
void require(bool cond, const char *msg);
// block offset of time row t of the G^R / G^< storage (row t of G^R holds
// G^R(t, 0..t), of G^< holds G^<(0..t, t)): t(t+1)/2, or its ring slot
inline size_t g_row(const Ctx &c, int t) { return c.gring ? (size_t)(t % c.gring) * (size_t)(c.nt + 1) : tri(t); }
// with ring storage of G, rows t0..t1 must be held (rows g_top - gring + 1 ..
// g_top); a no-op for the whole triangle
inline void g_need(const Ctx &c, int t0, int t1, const char *what) {
  if (!c.gring) return;
  t0 = t0 < 0 ? 0 : t0;
  if (t1 > c.g_top || t0 <= c.g_top - c.gring) {
    char msg[256];
    snprintf(msg, sizeof msg, "%s: needs G rows %d..%d, but the ring storage holds rows %d..%d", what, t0, t1,
             c.g_top - c.gring + 1 < 0 ? 0 : c.g_top - c.gring + 1, c.g_top);
    throw std::invalid_argument(msg);
  }
}
inline void g_wrote(Ctx &c, int t) { c.g_top = t > c.g_top ? t : c.g_top; }
// bytes of one stored Sigma^R / Sigma^< element
inline size_t s_elem(const Ctx &c) { return c.s32 ? 2 * sizeof(float) : 2 * sizeof(double); }

// Sigma: molGF2SolverDecomp::solve (src/mol2bSolver_decomp_impl.h:412-429)
void sigma_gf2(Ctx &c, int tstp, Prec prec);
// the same with no device allocation: the chunk scratch is a static device
// array of sigma_scratch_bytes() per GPU; sigma_scratch_init(device), with
// that device current, looks up its address once per device
void sigma_gf2_noalloc(Ctx &c, int tstp, Prec prec);
// the same slice by the full (non-decomposed) solver, molGF2Solver::solve
// (src/mol2bSolver_impl.h:5-120)
void sigma_gf2_full(Ctx &c, int tstp, Prec prec);
void sigma_scratch_init(int device);
size_t sigma_scratch_bytes();
void build_coulomb(Ctx &c, const double *Bq_host);
void g0_from_h0(Ctx &c, int tmax, const double *fock, double mu, double beta, double dt);
void tti_extend(Ctx &c, int which, int tsrc);

}  // namespace ned
