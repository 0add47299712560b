// Equilibrium (Matsubara) GF2 solver on the Gauss-Lobatto-Legendre grid: the
// state the reference takes from gfmol before propagating
// (src/molNEsim_decomp_impl.h:14-52 constructs gfmol::DecompSimulation,
// :50-52 runs it, :145-165 samples its G^M and Sigma^M on the collocation
// nodes; src/molNEsim.cpp:101-133 is the order of the calls).
//
// Unknown: G^M(tau_m) on the np = ntau+1 nodes, real for real orbitals.
// Equation (fermions, sig = -1):
//   G = G0 + G0 o Sigma o G,  (A o B)(tau) = int_0^beta A(tau - t) B(t) dt,
// with A(s < 0) = sig A(s + beta), G0 the Matsubara function of the Fock
// matrix F[rho] = h + 2J - K, rho = -G(beta), and Sigma the second-Born
// self-energy.  Sigma is evaluated by the real-time kernels on the T = 0
// slice: with G^tv(0, tau_m) = -i G^M(beta - tau_m) (the equilibrium form of
// the tv component), Sigma^M(beta - tau_m) = i Sigma^tv(0, tau_m).  This pairing
// of conventions is the one under which the tv Dyson equation
// (src/dyson_step.h:130-169) is stationary at t = 0.
//
// Discretisation: the convolution of nodal interpolants is exact
// (GLL::conv_weights).  In the eigenbasis of F, G0 is diagonal, so its
// convolution operator is block diagonal (one np x np block C_p per orbital)
// and the linear system
//   (I - C K_Sigma) G' = G0'
// (D = np * nao unknowns per column, rows ordered (m, orbital)) is assembled
// with an np*D*D product and solved by a blocked LU with partial pivoting on
// the GPU (panel: one workgroup; trailing update: the FP64 MFMA GEMM).
//
// Self-consistency: for fixed Sigma and F the chemical potential is found by
// a bracketed Illinois iteration on 2 tr rho(mu) = nelec (each evaluation one
// solve); the outer loop mixes G with Anderson acceleration.
#include "dyson.h"
#include "quadrature.h"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace ned {

namespace {

constexpr int LU_NB = 64;   // panel width (one wave64 lane per panel column)

inline unsigned nb256(long long n) { return (unsigned)((n + 255) / 256); }

// T = 0 slice of G from the real G^M: tv(0,m) = -i G^M(tau_{ntau-m}),
// les(0,0) = tv(0,0), ret(0,0) = -i
__global__ void kern_fill_t0(const double *Gr, int n, int np, z2 *ret0, z2 *les0, z2 *tv0) {
  const long long es = (long long)n * n;
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)np * es) return;
  const int m = (int)(idx / es), e = (int)(idx % es);
  const double g = Gr[(long long)(np - 1 - m) * es + e];
  tv0[idx] = z2{0.0, -g};
  if (m == 0) {
    les0[e] = z2{0.0, -g};
    ret0[e] = z2{0.0, (e / n == e % n) ? -1.0 : 0.0};
  }
}

// Sigma^M(tau_{ntau-m}) = i Sigma^tv(0, tau_m): real part into Sr, full value into Smat
__global__ void kern_sigma_mat(const z2 *Stv0, int n, int np, double *Sr, z2 *Smat) {
  const long long es = (long long)n * n;
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)np * es) return;
  const int m = (int)(idx / es), e = (int)(idx % es);
  const z2 s = Stv0[idx];
  const long long o = (long long)(np - 1 - m) * es + e;
  Sr[o] = -s.y;
  if (Smat) Smat[o] = z2{-s.y, s.x};
}

// G^M (real) -> mat blocks (complex)
__global__ void kern_store_mat(const double *Gr, long long cnt, z2 *mat) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < cnt) mat[idx] = z2{Gr[idx], 0.0};
}

// right-hand side G0' (rows (m,p), columns q): g[p][m] on the diagonal
__global__ void kern_rhs(const double *gT, int n, int np, double *Aug, long long lda, int D) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)D * n) return;
  const int row = (int)(idx / n), q = (int)(idx % n);
  const int m = row / n, p = row % n;
  Aug[(long long)row * lda + D + q] = (p == q) ? gT[(long long)p * np + m] : 0.0;
}

__global__ void kern_add_identity(double *A, long long lda, int D) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < D) A[(long long)i * lda + i] += 1.0;
}

// ---------------------------------------------------------------- LU
// One elimination step of the panel c0..c0+kb-1, spread over the whole chip.
// Step c >= 0 eliminates column col = c0+c from the rows below it with the
// pivot row already in prow (and stores the multipliers), and tracks the
// largest |entry| of column col+1 over those rows; step c = -1 only tracks
// column c0 over rows c0..D-1.  Each workgroup leaves its candidate in
// (pval, pidx); the last workgroup to finish (atomic ticket, fences on both
// sides) picks the pivot of the next column, swaps the two rows within the
// panel, records ipiv and loads the new pivot row into prow.  Rows map to
// waves, panel columns to lanes, so each row segment is one 512-byte access.
constexpr int LU_ROWS = 32;  // rows per workgroup and step
__global__ void __launch_bounds__(256) kern_lu_step(double *A, long long lda, int D, int c0, int kb, int c, double *prow,
                                                  double *pval, int *pidx, unsigned *ticket, int *ipiv, int *info) {
  __shared__ double sv[4];
  __shared__ int si[4];
  __shared__ bool last;
  const int tid = threadIdx.x, j = tid & 63, wv = tid >> 6;
  const int col = c0 + c, nxt = c + 1;  // nxt: panel index of the column searched
  const int r0 = col + 1 + blockIdx.x * LU_ROWS;
  const double pc = c >= 0 ? prow[c] : 1.0;
  const double pj = (c >= 0 && j < kb) ? prow[j] : 0.0;
  const double inv = 1.0 / pc;
  double best = -1.0;
  int bi = 0x7fffffff;
  for (int r = r0 + wv; r < min(D, r0 + LU_ROWS); r += 4) {
    double *row = A + (long long)r * lda + c0;
    double a = j < kb ? row[j] : 0.0;
    if (c >= 0) {
      const double l = __shfl(a, c, 64) * inv;
      if (j == c) a = l;
      else if (j > c) a -= l * pj;
      if (j >= c && j < kb) row[j] = a;
    }
    if (j == nxt && nxt < kb) {
      const double v = fabs(a);
      if (v > best) { best = v; bi = r; }
    }
  }
  // candidate of this wave sits in lane nxt; workgroup reduction
  if (nxt < kb) {
    const double bv = __shfl(best, nxt, 64);
    const int bx = __shfl(bi, nxt, 64);
    if (j == 0) { sv[wv] = bv; si[wv] = bx; }
  }
  __syncthreads();
  if (tid == 0) {
    double v = -1.0;
    int x = 0x7fffffff;
    if (nxt < kb)
      for (int h = 0; h < 4; h++)
        if (sv[h] > v || (sv[h] == v && si[h] < x)) { v = sv[h]; x = si[h]; }
    pval[blockIdx.x] = v;
    pidx[blockIdx.x] = x;
    __threadfence();
    last = atomicAdd(ticket, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  if (nxt >= kb) {
    if (tid == 0) *ticket = 0;
    return;
  }
  // pick the pivot of column c0+nxt over all workgroups
  if (tid < 64) {
    double v = -1.0;
    int x = 0x7fffffff;
    for (int b = tid; b < (int)gridDim.x; b += 64) {
      const double o = pval[b];
      const int oi = pidx[b];
      if (o > v || (o == v && oi < x)) { v = o; x = oi; }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(v, off);
      const int oi = __shfl_xor(x, off);
      if (o > v || (o == v && oi < x)) { v = o; x = oi; }
    }
    if (tid == 0) { sv[0] = v; si[0] = x; }
  }
  __syncthreads();
  const int pr = c0 + nxt, p = si[0];
  if (tid < kb) {
    double *a = A + (long long)pr * lda + c0 + tid, *b = A + (long long)p * lda + c0 + tid;
    const double ta = *a, tb = *b;
    if (p != pr) { *a = tb; *b = ta; }
    prow[tid] = p != pr ? tb : ta;
    if (tid == nxt && (p != pr ? tb : ta) == 0.0 && info[0] == 0) info[0] = pr + 1;
  }
  if (tid == 0) {
    ipiv[pr] = p;
    *ticket = 0;
  }
}

// apply the panel's row interchanges to columns j0..ncol-1
__global__ void kern_lu_swap(double *A, long long lda, int c0, int kb, const int *ipiv, int j0, int ncol) {
  const int j = j0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncol) return;
  for (int c = 0; c < kb; c++) {
    const int p = ipiv[c0 + c];
    if (p != c0 + c) {
      double *a = A + (long long)(c0 + c) * lda + j, *b = A + (long long)p * lda + j;
      const double t = *a;
      *a = *b;
      *b = t;
    }
  }
}

// Inverse of the kb x kb diagonal block at (i0, i0): unit lower (upper = 0)
// or upper triangular (upper = 1), into X (row-major, ld LU_NB).  One wave:
// lane c forms column c in LDS.
__global__ void __launch_bounds__(64) kern_tri_inv(const double *A, long long lda, int i0, int kb, int upper, double *X) {
  __shared__ double T[LU_NB][LU_NB + 1];
  __shared__ double Y[LU_NB][LU_NB + 1];
  const int c = threadIdx.x;
  for (int e = c; e < kb * kb; e += 64) T[e / kb][e % kb] = A[(long long)(i0 + e / kb) * lda + i0 + e % kb];
  __syncthreads();
  if (c < kb) {
    if (!upper) {
      for (int r = 0; r < kb; r++) {
        double v = (r == c) ? 1.0 : 0.0;
        if (r > c) {
          v = 0.0;
          for (int k = c; k < r; k++) v -= T[r][k] * Y[k][c];
        }
        Y[r][c] = v;
      }
    } else {
      for (int r = kb - 1; r >= 0; r--) {
        double v = 0.0;
        if (r <= c) {
          v = (r == c) ? 1.0 : 0.0;
          for (int k = r + 1; k <= c; k++) v -= T[r][k] * Y[k][c];
          v /= T[r][r];
        }
        Y[r][c] = v;
      }
    }
  }
  __syncthreads();
  for (int e = c; e < kb * kb; e += 64) X[(e / kb) * LU_NB + e % kb] = Y[e / kb][e % kb];
}

// ---------------------------------------------------------------- Anderson
// out[i] += <r, H_i> for i < nh (H rows of length L), out[nh] = max |r| (bits)
__global__ void kern_dots(const double *r, const double *H, int nh, long long L, double *out,
                          unsigned long long *amax) {
  __shared__ double red[8][256];
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double mx = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < L; i += (long long)gridDim.x * blockDim.x) {
    const double v = r[i];
    mx = fmax(mx, fabs(v));
#pragma unroll
    for (int h = 0; h < 8; h++)
      if (h < nh) acc[h] += v * H[(long long)h * L + i];
  }
  for (int h = 0; h < nh; h++) red[h][threadIdx.x] = acc[h];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int h = 0; h < nh; h++) red[h][threadIdx.x] += red[h][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < nh) atomicAdd(&out[threadIdx.x], red[threadIdx.x][0]);
  // max |r|: non-negative doubles order like their bit patterns
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
  if ((threadIdx.x & 63) == 0) atomicMax(amax, (unsigned long long)__double_as_longlong(mx));
}

// x = sum_i c_i v_i over up to 32 vectors
constexpr int COMBO_MAX = 32;
struct Combo {
  const double *v[COMBO_MAX];
  double c[COMBO_MAX];
  int nv;
};
__global__ void kern_combo(Combo k, long long L, double *x) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < L; i += (long long)gridDim.x * blockDim.x) {
    double s = 0;
    for (int j = 0; j < k.nv; j++) s += k.c[j] * k.v[j][i];
    x[i] = s;
  }
}

__global__ void kern_sub(const double *a, const double *b, long long L, double *out) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < L) out[i] = a[i] - b[i];
}

// real GEMM on row-major operands: C = alpha op(A) op(B) (+ C)
void dgemm(hipStream_t s, int M, int N, int K, double alpha, const double *A, long long ams, long long aks,
           const double *B, long long bks, long long bms, double *C, long long crs, int beta, int batch = 1,
           long long abs_ = 0, long long bbs = 0, long long cbs = 0) {
  GemmDesc d;
  d.M = M; d.N = N; d.K = K; d.batch = batch;
  d.A.ptr = A; d.A.ms = ams; d.A.ks_lo = aks; d.A.bs_lo = abs_;
  d.B.ptr = B; d.B.ms = bms; d.B.ks_lo = bks; d.B.bs_lo = bbs;
  d.C.ptr = C; d.C.rs = crs; d.C.cs = 1; d.C.bs_lo = cbs;
  d.alpha = alpha;
  d.beta = beta;
  gemm(d, PREC_F64, s);
}

}  // namespace

// Solve Aug[:, 0:D] X = Aug[:, D:D+nrhs] in place (X overwrites the
// right-hand sides); Aug is row-major D x (D + nrhs).  Returns the first zero
// pivot (1-based) or 0.
int lu_solve_aug(hipStream_t s, double *Aug, int D, int nrhs, int *ipiv, int *info, const LuWork &ws) {
  const long long lda = (long long)D + nrhs;
  const int ncol = D + nrhs;
  NED_HIP(hipMemsetAsync(info, 0, sizeof(int), s));
  NED_HIP(hipMemsetAsync(ws.ticket, 0, sizeof(unsigned), s));
  for (int c0 = 0; c0 < D; c0 += LU_NB) {
    const int kb = std::min(LU_NB, D - c0);
    for (int c = -1; c < kb; c++) {
      const int rows = D - (c0 + c + 1);
      const unsigned grid = (unsigned)std::max(1, (rows + LU_ROWS - 1) / LU_ROWS);
      hipLaunchKernelGGL(kern_lu_step, dim3(grid), dim3(256), 0, s, Aug, lda, D, c0, kb, c, ws.prow, ws.pval, ws.pidx,
                         ws.ticket, ipiv, info);
    }
    const int j0 = c0 + kb, nc = ncol - j0;
    if (nc <= 0) continue;
    hipLaunchKernelGGL(kern_lu_swap, dim3(nb256(nc)), dim3(256), 0, s, Aug, lda, c0, kb, ipiv, j0, ncol);
    // U12 = L11^{-1} A12 as a GEMM with the explicit inverse
    hipLaunchKernelGGL(kern_tri_inv, dim3(1), dim3(64), 0, s, Aug, lda, c0, kb, 0, ws.tinv);
    NED_HIP(hipMemcpy2DAsync(ws.tmp, (size_t)nc * sizeof(double), Aug + (long long)c0 * lda + j0, (size_t)lda * sizeof(double),
                             (size_t)nc * sizeof(double), kb, hipMemcpyDeviceToDevice, s));
    dgemm(s, kb, nc, kb, 1.0, ws.tinv, LU_NB, 1, ws.tmp, nc, 1, Aug + (long long)c0 * lda + j0, lda, 0);
    const int mr = D - j0;
    if (mr > 0)
      dgemm(s, mr, nc, kb, -1.0, Aug + (long long)j0 * lda + c0, lda, 1, Aug + (long long)c0 * lda + j0, lda, 1,
            Aug + (long long)j0 * lda + j0, lda, 1);
  }
  for (int i0 = ((D - 1) / LU_NB) * LU_NB; i0 >= 0; i0 -= LU_NB) {
    const int kb = std::min(LU_NB, D - i0), kr = D - i0 - kb;
    if (kr > 0)
      dgemm(s, kb, nrhs, kr, -1.0, Aug + (long long)i0 * lda + i0 + kb, lda, 1, Aug + (long long)(i0 + kb) * lda + D,
            lda, 1, Aug + (long long)i0 * lda + D, lda, 1);
    hipLaunchKernelGGL(kern_tri_inv, dim3(1), dim3(64), 0, s, Aug, lda, i0, kb, 1, ws.tinv);
    NED_HIP(hipMemcpy2DAsync(ws.tmp, (size_t)nrhs * sizeof(double), Aug + (long long)i0 * lda + D, (size_t)lda * sizeof(double),
                             (size_t)nrhs * sizeof(double), kb, hipMemcpyDeviceToDevice, s));
    dgemm(s, kb, nrhs, kb, 1.0, ws.tinv, LU_NB, 1, ws.tmp, nrhs, 1, Aug + (long long)i0 * lda + D, lda, 0);
  }
  check_launch("lu_solve");
  int h = 0;
  NED_HIP(hipMemcpyAsync(&h, info, sizeof(int), hipMemcpyDeviceToHost, s));
  NED_HIP(hipStreamSynchronize(s));
  return h;
}

namespace {

double fermi(double beta, double w) {
  const double a = w * beta;
  if (a > 700) return 0.0;
  if (a < -700) return 1.0;
  return 1.0 / (1.0 + std::exp(a));
}

// -G0 of one level: eigenvalue eps of F, at tau in [0, beta]
double g0_level(double eps, double mu, double tau, double beta) {
  const double l = mu - eps;
  return l < 0 ? -std::exp(l * tau) * fermi(beta, l) : -std::exp((tau - beta) * l) * fermi(beta, -l);
}

struct MatSolver {
  Ctx &c;
  int n, np, D;
  long long es, L;
  double beta;
  std::vector<double> tau;
  DevBuf w, Gr, Gn, Sr, Sp, tmp, K, Aug, gT, Cp, ipiv, info, V, dots, luw;
  LuWork ws{};
  std::vector<double> eps, Vh;
  int nsolve = 0;

  MatSolver(Ctx &cc, double b) : c(cc), n(cc.nao), np(cc.np), beta(b) {
    D = np * n;
    es = (long long)n * n;
    L = (long long)np * es;
    GLL g(c.ntau);
    std::vector<double> wh;
    g.conv_weights(beta, c.sig, wh);
    w.ensure(wh.size() * sizeof(double));
    NED_HIP(hipMemcpyAsync(w.p, wh.data(), wh.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    tau.resize(np);
    for (int m = 0; m < np; m++) tau[m] = beta * (g.x[m] + 1) / 2;
    for (DevBuf *b2 : {&Gr, &Gn, &Sr, &Sp, &tmp}) b2->ensure(L * sizeof(double));
    K.ensure((size_t)D * D * sizeof(double));
    Aug.ensure((size_t)D * (D + n) * sizeof(double));
    gT.ensure((size_t)n * np * sizeof(double));
    Cp.ensure((size_t)n * np * np * sizeof(double));
    ipiv.ensure((size_t)D * sizeof(int));
    info.ensure(sizeof(int));
    const size_t nwg = (size_t)D / 32 + 2;
    const size_t ntmp = (size_t)LU_NB * (D + n);
    luw.ensure((LU_NB + LU_NB * LU_NB + ntmp + nwg) * sizeof(double) + (nwg + 1) * sizeof(int));
    ws.prow = luw.as<double>();
    ws.tinv = ws.prow + LU_NB;
    ws.tmp = ws.tinv + LU_NB * LU_NB;
    ws.pval = ws.tmp + ntmp;
    ws.pidx = reinterpret_cast<int *>(ws.pval + nwg);
    ws.ticket = reinterpret_cast<unsigned *>(ws.pidx + nwg);
    V.ensure(es * sizeof(double));
    dots.ensure(32 * sizeof(double));
    NED_HIP(hipStreamSynchronize(c.stream));  // wh is released on return
  }

  // Sigma^M[G] (real) into Sr, F[G] on the host; eigenbasis of F; Sigma' = V^T Sigma V; K_Sigma'
  void prepare(const double *h1, bool corr, std::vector<double> &F) {
    hipStream_t s = c.stream;
    hipLaunchKernelGGL(kern_fill_t0, dim3(nb256(L)), dim3(256), 0, s, Gr.as<double>(), n, np, c.G.ret.as<z2>(),
                       c.G.les.as<z2>(), c.G.tv.as<z2>());
    check_launch("matsubara: T=0 slice");
    hf_decomp(c, 0, h1);
    F.resize(es);
    std::vector<z2> Fz(es);
    NED_HIP(hipMemcpyAsync(Fz.data(), c.hmf.p, es * sizeof(z2), hipMemcpyDeviceToHost, s));
    NED_HIP(hipStreamSynchronize(s));
    for (long long e = 0; e < es; e++) F[e] = 0.5 * (Fz[e].x + Fz[(e % n) * n + e / n].x);
    eps.resize(n);
    Vh.resize(es);
    sym_eig(n, F.data(), eps.data(), Vh.data());
    NED_HIP(hipMemcpyAsync(V.p, Vh.data(), es * sizeof(double), hipMemcpyHostToDevice, s));
    if (corr) {
      sigma_gf2(c, 0, PREC_F64);
      hipLaunchKernelGGL(kern_sigma_mat, dim3(nb256(L)), dim3(256), 0, s, c.S.tv.as<z2>(), n, np, Sr.as<double>(),
                         c.S.mat.p ? c.S.mat.as<z2>() : nullptr);
      check_launch("matsubara: Sigma^M");
    } else {
      NED_HIP(hipMemsetAsync(Sr.p, 0, L * sizeof(double), s));
    }
    // Sigma'_j = V^T Sigma_j V
    dgemm(s, n, n, n, 1.0, Sr.as<double>(), n, 1, V.as<double>(), n, 1, tmp.as<double>(), n, 0, np, es, 0, es);
    dgemm(s, n, n, n, 1.0, V.as<double>(), 1, n, tmp.as<double>(), n, 1, Sp.as<double>(), n, 0, np, 0, es, es);
    // K[(m,p),(m',q)] = sum_j w[m][j][m'] Sigma'_j[p][q]; batch (m, p): rows m', columns q
    GemmDesc d;
    d.M = np; d.N = n; d.K = np; d.batch = np * n;
    d.A.ptr = w.p; d.A.ms = 1; d.A.ks_lo = np; d.A.bdiv = n; d.A.bs_hi = (long long)np * np; d.A.bs_lo = 0;
    d.B.ptr = Sp.p; d.B.ms = 1; d.B.ks_lo = es; d.B.bdiv = n; d.B.bs_hi = 0; d.B.bs_lo = n;
    d.C.ptr = K.p; d.C.rs = n; d.C.cs = 1; d.C.bs_lo = D;
    gemm(d, PREC_F64, s);
  }

  // G' for the current K and mu into Aug's right-hand-side columns; returns 2 tr rho
  double solve(double mu) {
    hipStream_t s = c.stream;
    nsolve++;
    std::vector<double> g((size_t)n * np);
    for (int p = 0; p < n; p++)
      for (int m = 0; m < np; m++) g[(size_t)p * np + m] = g0_level(eps[p], mu, tau[m], beta);
    NED_HIP(hipMemcpyAsync(gT.p, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice, s));
    // C_p[m][m'] = sum_j w[m][j][m'] g_p(tau_j): batch m, rows p, columns m'
    dgemm(s, n, np, np, 1.0, gT.as<double>(), np, 1, w.as<double>(), np, 1, Cp.as<double>(), (long long)np * np, 0, np,
          0, (long long)np * np, np);
    // Aug[(m,p), 0:D] = -sum_m'' C_p[m][m''] K[(m'',p), :]: batch p
    const long long lda = (long long)D + n;
    dgemm(s, np, D, np, -1.0, Cp.as<double>(), np, 1, K.as<double>(), (long long)n * D, 1, Aug.as<double>(), n * lda, 0,
          n, (long long)np * np, D, lda);
    hipLaunchKernelGGL(kern_add_identity, dim3(nb256(D)), dim3(256), 0, s, Aug.as<double>(), lda, D);
    hipLaunchKernelGGL(kern_rhs, dim3(nb256((long long)D * n)), dim3(256), 0, s, gT.as<double>(), n, np,
                       Aug.as<double>(), lda, D);
    check_launch("matsubara: system");
    int bad = lu_solve_aug(s, Aug.as<double>(), D, n, ipiv.as<int>(), info.as<int>(), ws);
    if (bad) throw std::runtime_error("matsubara: singular Dyson system (the iteration did not converge)");
    // 2 tr G'(beta) rows (ntau, p), column p
    std::vector<double> last((size_t)n * lda);
    NED_HIP(hipMemcpyAsync(last.data(), Aug.as<double>() + (size_t)(np - 1) * n * lda, last.size() * sizeof(double),
                           hipMemcpyDeviceToHost, s));
    NED_HIP(hipStreamSynchronize(s));
    double tr = 0;
    for (int p = 0; p < n; p++) tr += last[(size_t)p * lda + D + p];
    return -2.0 * tr;
  }

  // Gn = V G' V^T from Aug's right-hand sides
  void rotate_back() {
    hipStream_t s = c.stream;
    const long long lda = (long long)D + n;
    // tmp_m = G'_m V^T: G'_m rows p at Aug[(m*n+p)*lda + D + q]
    dgemm(s, n, n, n, 1.0, Aug.as<double>() + D, lda, 1, V.as<double>(), 1, n, tmp.as<double>(), n, 0, np, n * lda, 0,
          es);
    dgemm(s, n, n, n, 1.0, V.as<double>(), n, 1, tmp.as<double>(), n, 1, Gn.as<double>(), n, 0, np, 0, es, es);
  }
};

}  // namespace

void matsubara_gf2(Ctx &c, const double *h1, double nelec, double beta, bool corr, bool fix_mu, double *mu_io,
                   int maxiter, double tol, double *fock_out, int *iters_out, double *err_out) {
  require(c.have_coulomb, "matsubara: Coulomb tensor not set");
  require(beta > 0, "matsubara: beta must be positive");
  require(c.sig == -1, "matsubara: fermions only");
  require(fix_mu || nelec > 0, "matsubara: nelec must be positive");
  hipStream_t s = c.stream;
  MatSolver ms(c, beta);
  const int n = c.nao, np = c.np;
  const long long L = ms.L;
  std::vector<double> F(ms.es);
  double mu = *mu_io;

  // start: G0 of the Fock matrix of the given G^M, or of h1 if G^M is empty
  {
    std::vector<z2> gm(L);
    NED_HIP(hipMemcpyAsync(gm.data(), c.G.mat.p, L * sizeof(z2), hipMemcpyDeviceToHost, s));
    NED_HIP(hipStreamSynchronize(s));
    std::vector<double> gr(L);
    double nz = 0;
    for (long long i = 0; i < L; i++) { gr[i] = gm[i].x; nz = std::max(nz, std::fabs(gm[i].x)); }
    if (nz == 0) {
      // G^M of h1 with mu placing nelec/2 levels below it
      std::vector<double> e(n), Vv(ms.es);
      sym_eig(n, h1, e.data(), Vv.data());
      if (!fix_mu) {
        const int occ = std::max(1, std::min(n - 1, (int)std::lround(nelec / 2)));
        mu = 0.5 * (e[occ - 1] + e[std::min(occ, n - 1)]);
      }
      for (int m = 0; m < np; m++)
        for (int a = 0; a < n; a++)
          for (int b = 0; b < n; b++) {
            double v = 0;
            for (int p = 0; p < n; p++) v += Vv[a * n + p] * g0_level(e[p], mu, ms.tau[m], beta) * Vv[b * n + p];
            gr[(size_t)m * ms.es + a * n + b] = v;
          }
    }
    NED_HIP(hipMemcpyAsync(ms.Gr.p, gr.data(), L * sizeof(double), hipMemcpyHostToDevice, s));
    NED_HIP(hipStreamSynchronize(s));
  }

  // Anderson history (type II): X_i, R_i
  constexpr int HIST = 6;
  DevBuf Xh, Rh, r;
  Xh.ensure((size_t)HIST * L * sizeof(double));
  Rh.ensure((size_t)HIST * L * sizeof(double));
  r.ensure(L * sizeof(double));
  int nh = 0, head = 0;
  std::vector<double> RR(HIST * HIST, 0.0);  // <R_i, R_j> of the stored residuals
  double err = INFINITY;
  double dmu = 0.5, slope = 0.0;
  const bool verbose = std::getenv("NED_MAT_VERBOSE") != nullptr;
  int it = 0;
  for (; it < maxiter; it++) {
    ms.prepare(h1, corr, F);
    if (fix_mu) {
      ms.solve(mu);
    } else {
      // 2 tr rho(mu) increases with mu: secant from the previous mu (with the
      // previous slope), extended until the root is bracketed, then Illinois.
      // The last solve is always at the returned mu (rotate_back uses it).
      const double ftol = 1e-12 * std::max(1.0, nelec);
      double last_x = NAN;
      auto f = [&](double m) {
        last_x = m;
        return ms.solve(m) - nelec;
      };
      double x0 = mu, f0 = f(x0), x = x0, fx = f0;
      if (std::fabs(f0) > ftol) {
        double st = slope > 0 ? -f0 / slope : (f0 < 0 ? 1.0 : -1.0) * std::max(1e-3, 2 * dmu);
        st = std::copysign(std::min(std::max(std::fabs(st), 1e-10), 2.0), f0 < 0 ? 1.0 : -1.0);
        double x1 = x0 + st, f1 = f(x1);
        int e = 0;
        while (f0 * f1 > 0 && std::fabs(f1) > ftol && e++ < 60) {
          // same side: secant step if it moves the right way, else double the step
          const double dir = f1 < 0 ? 1.0 : -1.0;
          double sx = (f1 != f0) ? -f1 * (x1 - x0) / (f1 - f0) : 0.0;
          const double lim = 4 * std::fabs(x1 - x0);
          if (!(sx * dir > 0)) sx = dir * 2 * std::fabs(x1 - x0);
          if (std::fabs(sx) > lim) sx = dir * lim;
          x0 = x1; f0 = f1;
          x1 = x0 + sx;
          f1 = f(x1);
        }
        if (f0 * f1 > 0 && std::fabs(f1) > ftol)
          throw std::runtime_error("matsubara: the particle number did not converge (no bracketing mu)");
        if (f1 != f0) slope = (f1 - f0) / (x1 - x0);
        x = x1; fx = f1;
        if (std::fabs(f1) > ftol) {
          double a = x0, fa = f0, b = x1, fb = f1;
          if (fa > 0) { std::swap(a, b); std::swap(fa, fb); }
          int side = 0;
          for (int it2 = 0; it2 < 200; it2++) {
            x = (a * fb - b * fa) / (fb - fa);
            fx = f(x);
            if (std::fabs(fx) < ftol || std::fabs(b - a) < 1e-14 * std::max(1.0, std::fabs(x))) break;
            if (fx < 0) {
              a = x; fa = fx;
              if (side == -1) fb *= 0.5;
              side = -1;
            } else {
              b = x; fb = fx;
              if (side == 1) fa *= 0.5;
              side = 1;
            }
          }
        }
      }
      if (last_x != x) f(x);
      dmu = std::fabs(x - mu);
      mu = x;
    }
    ms.rotate_back();
    // r = Gn - G; <r, R_i>; max |r|
    hipLaunchKernelGGL(kern_sub, dim3(nb256(L)), dim3(256), 0, s, ms.Gn.as<double>(), ms.Gr.as<double>(), L, r.as<double>());
    NED_HIP(hipMemsetAsync(ms.dots.p, 0, 32 * sizeof(double), s));
    unsigned long long *amax = reinterpret_cast<unsigned long long *>(ms.dots.as<double>() + 16);
    hipLaunchKernelGGL(kern_dots, dim3(256), dim3(256), 0, s, r.as<double>(), Rh.as<double>(), nh, L,
                       ms.dots.as<double>(), amax);
    // <r, r> with a second pass (one vector)
    hipLaunchKernelGGL(kern_dots, dim3(256), dim3(256), 0, s, r.as<double>(), r.as<double>(), 1, L,
                       ms.dots.as<double>() + 8, amax);
    check_launch("matsubara: residual");
    double hd[17];
    NED_HIP(hipMemcpyAsync(hd, ms.dots.p, sizeof(hd), hipMemcpyDeviceToHost, s));
    NED_HIP(hipStreamSynchronize(s));
    unsigned long long bits;
    std::memcpy(&bits, &hd[16], sizeof(bits));
    double em;
    std::memcpy(&em, &bits, sizeof(em));
    err = em;
    if (verbose) fprintf(stderr, "[matsubara] iteration %d: max|dG| %.3e mu %.12f solves %d\n", it, err, mu, ms.nsolve);
    if (!std::isfinite(err)) throw std::runtime_error("matsubara: non-finite update");
    if (err < tol) {
      NED_HIP(hipMemcpyAsync(ms.Gr.p, ms.Gn.p, L * sizeof(double), hipMemcpyDeviceToDevice, s));
      it++;
      break;
    }
    // order of the stored entries, oldest first
    std::vector<int> ord;
    for (int i = 0; i < nh; i++) ord.push_back((head - nh + i + HIST) % HIST);
    // Anderson step over the differences of consecutive entries plus the current one
    Combo k{};
    k.nv = 0;
    static_assert(2 + 4 * HIST <= COMBO_MAX, "Anderson combination exceeds Combo");
    auto add = [&](const double *v, double cc) {
      if (k.nv >= COMBO_MAX) throw std::logic_error("matsubara: Anderson combination too long");
      k.v[k.nv] = v;
      k.c[k.nv] = cc;
      k.nv++;
    };
    add(ms.Gr.as<double>(), 1.0);
    add(r.as<double>(), 1.0);
    if (nh >= 1) {
      // dR_i = R_{i+1} - R_i (i = 0..nh-1, with R_nh = r), dX_i likewise with X_nh = G
      const int md = nh;
      auto rr = [&](int i, int j) -> double {  // <R_i, R_j> (logical order; index nh is the current r)
        if (i == nh && j == nh) return hd[8];
        if (i == nh) return hd[ord[j]];
        if (j == nh) return hd[ord[i]];
        return RR[ord[i] * HIST + ord[j]];
      };
      std::vector<double> A(md * md), bvec(md);
      for (int i = 0; i < md; i++) {
        for (int j = 0; j < md; j++)
          A[i * md + j] = rr(i + 1, j + 1) - rr(i + 1, j) - rr(i, j + 1) + rr(i, j);
        bvec[i] = rr(i + 1, nh) - rr(i, nh);
      }
      double tr = 0;
      for (int i = 0; i < md; i++) tr += A[i * md + i];
      for (int i = 0; i < md; i++) A[i * md + i] += 1e-12 * tr + 1e-300;
      // solve A gamma = b (Gaussian elimination with partial pivoting, md <= HIST)
      std::vector<double> gmm = bvec;
      for (int col = 0; col < md; col++) {
        int p = col;
        for (int rr2 = col + 1; rr2 < md; rr2++)
          if (std::fabs(A[rr2 * md + col]) > std::fabs(A[p * md + col])) p = rr2;
        if (p != col) {
          for (int j = 0; j < md; j++) std::swap(A[col * md + j], A[p * md + j]);
          std::swap(gmm[col], gmm[p]);
        }
        for (int rr2 = col + 1; rr2 < md; rr2++) {
          const double l = A[rr2 * md + col] / A[col * md + col];
          for (int j = col; j < md; j++) A[rr2 * md + j] -= l * A[col * md + j];
          gmm[rr2] -= l * gmm[col];
        }
      }
      for (int col = md - 1; col >= 0; col--) {
        double sacc = gmm[col];
        for (int j = col + 1; j < md; j++) sacc -= A[col * md + j] * gmm[j];
        gmm[col] = sacc / A[col * md + col];
      }
      // G_new = G + r - sum_i gamma_i (dX_i + dR_i); dX_i + dR_i = (X+R)_{i+1} - (X+R)_i
      for (int i = 0; i < md; i++) {
        const double gi = gmm[i];
        // entry i+1 (the current one when i+1 == nh)
        if (i + 1 == nh) {
          add(ms.Gr.as<double>(), -gi);
          add(r.as<double>(), -gi);
        } else {
          add(Xh.as<double>() + (size_t)ord[i + 1] * L, -gi);
          add(Rh.as<double>() + (size_t)ord[i + 1] * L, -gi);
        }
        add(Xh.as<double>() + (size_t)ord[i] * L, gi);
        add(Rh.as<double>() + (size_t)ord[i] * L, gi);
      }
    }
    // store the current (G, r) as the newest entry, then form the update in Gn
    hipLaunchKernelGGL(kern_combo, dim3(1024), dim3(256), 0, s, k, L, ms.Gn.as<double>());
    {
      // the slot about to be overwritten may be referenced by k: k ran first on the stream
      const int slot = head;
      NED_HIP(hipMemcpyAsync(Xh.as<double>() + (size_t)slot * L, ms.Gr.p, L * sizeof(double), hipMemcpyDeviceToDevice, s));
      NED_HIP(hipMemcpyAsync(Rh.as<double>() + (size_t)slot * L, r.p, L * sizeof(double), hipMemcpyDeviceToDevice, s));
      // Gram entries of the new slot: <r, R_i> were computed before the overwrite
      for (int i = 0; i < nh; i++) {
        const int o = ord[i];
        if (o == slot) continue;
        RR[slot * HIST + o] = RR[o * HIST + slot] = hd[o];
      }
      RR[slot * HIST + slot] = hd[8];
      head = (head + 1) % HIST;
      nh = std::min(nh + 1, HIST);
    }
    NED_HIP(hipMemcpyAsync(ms.Gr.p, ms.Gn.p, L * sizeof(double), hipMemcpyDeviceToDevice, s));
    check_launch("matsubara: Anderson");
  }
  *iters_out = it;
  *err_out = err;
  *mu_io = mu;
  // final state: G^M, and Sigma^M / F consistent with it
  hipLaunchKernelGGL(kern_store_mat, dim3(nb256(L)), dim3(256), 0, s, ms.Gr.as<double>(), L, c.G.mat.as<z2>());
  check_launch("matsubara: store");
  std::vector<double> Ff;
  ms.prepare(h1, corr, Ff);
  for (long long e = 0; e < ms.es; e++) fock_out[e] = Ff[e];
  NED_HIP(hipStreamSynchronize(s));
  invalidate_minv(c);
  c.have_mix = false;
}

}  // namespace ned
