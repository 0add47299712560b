// Decomposed second-Born self-energy and HF mean field, restating
// src/mol2bSolver_decomp_impl.h and src/molHFSolver.cpp.  Each time point is
// the reference's chain of GEMMs, in the reference's order, with the scratch
// tensors X1..X3, Y1, Y2, P1, P2, Z of src/mol2bSolver.h:80-88 held per call.
// (Test infrastructure only.)
#include "oracle.h"
#include <cstring>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

namespace {
struct Scratch {
  std::vector<cplx> X1, X2, X3, Y1, Y2, P1, P2, Z, tmp;
  Scratch(int n, int a)
      : X1((size_t)n * n * a), X2((size_t)n * n * a), X3((size_t)n * n * a), Y1((size_t)n * n * a),
        Y2((size_t)n * n * a), P1((size_t)a * a), P2((size_t)a * a), Z((size_t)n * n * n * n),
        tmp((size_t)n * n) {}
};

// out[b] (c x r) = in[b] (r x c)^T for nb blocks of r*c elements
void block_transpose(int nb, int r, int c, const cplx *in, cplx *out) {
  for (int b = 0; b < nb; b++) {
    const cplx *I = in + (size_t)b * r * c;
    cplx *O = out + (size_t)b * r * c;
    for (int i = 0; i < r; i++)
      for (int j = 0; j < c; j++) O[(size_t)j * r + i] = I[(size_t)i * c + j];
  }
}
const cplx ONE(1.0), ZERO(0.0);
}  // namespace

GF2Decomp::GF2Decomp(int n, int a, const double *V) : nao(n), naux(a), Vija(V), Viaj((size_t)n * n * a) {
  // tests/mol2bSolver-test.cpp:97-104
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int q = 0; q < a; q++) Viaj[((size_t)i * a + q) * n + j] = V[((size_t)i * n + j) * a + q];
}

// src/mol2bSolver_decomp_impl.h:6-49
static void bubble_les(const GF2Decomp &s, Scratch &w, int tstp, int t, const Green &Sg, const Green &G) {
  const int n = s.nao, a = s.naux, n2 = n * n, naa = n * a;
  const double *Viaj = s.Viaj.data(), *Vija = s.Vija;
  const cplx *L = G.lesptr(t, tstp), *R = G.retptr(tstp, t);
  gemm('N', 'N', naa, n, n, ONE, Viaj, L, ZERO, w.X1.data());
  block_transpose(n, a, n, w.X1.data(), w.Y1.data());
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) w.tmp[i * n + j] = R[j * n + i] - std::conj(L[i * n + j]);
  gemm('N', 'T', n, naa, n, ONE, w.tmp.data(), Viaj, ZERO, w.X1.data());
  gemm('T', 'N', a, a, n2, ONE, w.Y1.data(), w.X1.data(), ZERO, w.P1.data());
  gemm('N', 'T', n2, a, a, ONE, Vija, w.P1.data(), ZERO, w.X1.data());
  block_transpose(n, n, a, w.X1.data(), w.X2.data());
  gemm('N', 'N', n, n, naa, cplx(2.0), w.Y1.data(), w.X2.data(), ONE, Sg.lesptr(t, tstp));
}

// src/mol2bSolver_decomp_impl.h:51-91
static void bubble_tv(const GF2Decomp &s, Scratch &w, int tstp, int t, const Green &Sg, const Green &G) {
  const int n = s.nao, a = s.naux, n2 = n * n, naa = n * a, ntau = G.ntau, sig = G.sig;
  const double *Viaj = s.Viaj.data(), *Vija = s.Vija;
  gemm('N', 'N', naa, n, n, ONE, Viaj, G.tvptr(tstp, t), ZERO, w.X1.data());
  block_transpose(n, a, n, w.X1.data(), w.Y1.data());
  gemm('C', 'T', n, naa, n, cplx(-(double)sig), G.tvptr(tstp, ntau - t), Viaj, ZERO, w.X1.data());
  gemm('T', 'N', a, a, n2, ONE, w.Y1.data(), w.X1.data(), ZERO, w.P1.data());
  gemm('N', 'T', n2, a, a, ONE, Vija, w.P1.data(), ZERO, w.X1.data());
  block_transpose(n, n, a, w.X1.data(), w.X2.data());
  gemm('N', 'N', n, n, naa, cplx(2.0), w.Y1.data(), w.X2.data(), ONE, Sg.tvptr(tstp, t));
}

// src/mol2bSolver_decomp_impl.h:94-172
static void bubble_ret(const GF2Decomp &s, Scratch &w, int tstp, int t, const Green &Sg, const Green &G) {
  const int n = s.nao, a = s.naux, n2 = n * n, naa = n * a;
  const double *Viaj = s.Viaj.data(), *Vija = s.Vija;
  const cplx *L = G.lesptr(t, tstp), *R = G.retptr(tstp, t);
  gemm('N', 'H', naa, n, n, cplx(-1.0), Viaj, L, ZERO, w.X1.data());
  gemm('N', 'N', naa, n, n, ONE, Viaj, R, ZERO, w.X2.data());
  block_transpose(n, a, n, w.X1.data(), w.Y1.data());
  block_transpose(n, a, n, w.X2.data(), w.Y2.data());
  gemm('C', 'T', n, naa, n, ONE, R, Viaj, ZERO, w.X1.data());
  gemm('T', 'T', n, naa, n, ONE, L, Viaj, ZERO, w.X2.data());
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) w.tmp[i * n + j] = L[j * n + i] - std::conj(R[i * n + j]);
  gemm('N', 'T', n, naa, n, ONE, w.tmp.data(), Viaj, ZERO, w.X3.data());
  gemm('T', 'N', a, a, n2, ONE, w.Y1.data(), w.X3.data(), ZERO, w.P1.data());
  gemm('T', 'N', a, a, n2, ONE, w.Y2.data(), w.X2.data(), ZERO, w.P2.data());
  gemm('T', 'N', a, a, n2, ONE, w.Y1.data(), w.X1.data(), ONE, w.P2.data());
  gemm('N', 'T', n2, a, a, ONE, Vija, w.P1.data(), ZERO, w.X1.data());
  gemm('N', 'T', n2, a, a, ONE, Vija, w.P2.data(), ZERO, w.X2.data());
  for (int kk = 0; kk < n; kk++) {
    block_transpose(1, n, a, w.X2.data() + (size_t)kk * naa, w.X3.data() + (size_t)kk * naa);
    block_transpose(1, n, a, w.X1.data() + (size_t)kk * naa, w.X2.data() + (size_t)kk * naa);
  }
  cplx *S = Sg.retptr(tstp, t);
  std::vector<cplx> acc((size_t)n2, ZERO);
  gemm('N', 'N', n, n, naa, ONE, w.Y2.data(), w.X3.data(), ZERO, acc.data());
  gemm('N', 'N', n, n, naa, ONE, w.Y1.data(), w.X3.data(), ONE, acc.data());
  gemm('N', 'N', n, n, naa, ONE, w.Y2.data(), w.X2.data(), ONE, acc.data());
  for (int e = 0; e < n2; e++) S[e] += 2.0 * acc[e];
}

// src/mol2bSolver_decomp_impl.h:182-233
static void exch_les(const GF2Decomp &s, Scratch &w, int tstp, int t, const Green &Sg, const Green &G) {
  const int n = s.nao, a = s.naux, n2 = n * n, n3 = n2 * n, naa = n * a;
  const double *Viaj = s.Viaj.data(), *Vija = s.Vija;
  const cplx *L = G.lesptr(t, tstp), *R = G.retptr(tstp, t);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) w.tmp[i * n + j] = R[j * n + i] - std::conj(L[i * n + j]);
  gemm('N', 'T', n, naa, n, ONE, w.tmp.data(), Viaj, ZERO, w.X1.data());
  block_transpose(n, n, a, w.X1.data(), w.X2.data());
  gemm('N', 'T', naa, n, n, ONE, w.X2.data(), L, ZERO, w.X1.data());
  for (int nn = 0; nn < n; nn++)
    gemm('N', 'N', n2, n, a, ONE, Vija, w.X1.data() + (size_t)nn * naa, ZERO, w.Z.data() + (size_t)nn * n3);
  gemm('N', 'N', naa, n, n, ONE, Viaj, L, ZERO, w.Y1.data());
  block_transpose(n, a, n, w.Y1.data(), w.Y2.data());
  gemm('T', 'N', n2, a, n2, ONE, w.Z.data(), w.Y2.data(), ZERO, w.X1.data());
  gemm('N', 'T', n, n, naa, cplx(-1.0), Vija, w.X1.data(), ONE, Sg.lesptr(t, tstp));
}

// src/mol2bSolver_decomp_impl.h:236-289
static void exch_tv(const GF2Decomp &s, Scratch &w, int tstp, int t, const Green &Sg, const Green &G) {
  const int n = s.nao, a = s.naux, n2 = n * n, n3 = n2 * n, naa = n * a, ntau = G.ntau, sig = G.sig;
  const double *Viaj = s.Viaj.data(), *Vija = s.Vija;
  gemm('C', 'T', n, naa, n, cplx(-(double)sig), G.tvptr(tstp, ntau - t), Viaj, ZERO, w.X1.data());
  block_transpose(n, n, a, w.X1.data(), w.X2.data());
  gemm('N', 'T', naa, n, n, ONE, w.X2.data(), G.tvptr(tstp, t), ZERO, w.X1.data());
  for (int nn = 0; nn < n; nn++)
    gemm('N', 'N', n2, n, a, ONE, Vija, w.X1.data() + (size_t)nn * naa, ZERO, w.Z.data() + (size_t)nn * n3);
  gemm('N', 'N', naa, n, n, ONE, Viaj, G.tvptr(tstp, t), ZERO, w.Y1.data());
  block_transpose(n, a, n, w.Y1.data(), w.Y2.data());
  gemm('T', 'N', n2, a, n2, ONE, w.Z.data(), w.Y2.data(), ZERO, w.X1.data());
  gemm('N', 'T', n, n, naa, cplx(-1.0), Vija, w.X1.data(), ONE, Sg.tvptr(tstp, t));
}

// src/mol2bSolver_decomp_impl.h:293-402
static void exch_ret(const GF2Decomp &s, Scratch &w, int tstp, int t, const Green &Sg, const Green &G) {
  const int n = s.nao, a = s.naux, n2 = n * n, n3 = n2 * n, naa = n * a;
  const double *Viaj = s.Viaj.data(), *Vija = s.Vija;
  const cplx *L = G.lesptr(t, tstp), *R = G.retptr(tstp, t);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) w.tmp[i * n + j] = -std::conj(R[i * n + j]) + L[j * n + i];
  gemm('N', 'T', n, naa, n, ONE, w.tmp.data(), Viaj, ZERO, w.X2.data());
  block_transpose(n, n, a, w.X2.data(), w.Y1.data());
  gemm('T', 'T', n, naa, n, ONE, L, Viaj, ZERO, w.X2.data());
  block_transpose(n, n, a, w.X2.data(), w.Y2.data());
  gemm('C', 'T', n, naa, n, ONE, R, Viaj, ZERO, w.X2.data());
  block_transpose(n, n, a, w.X2.data(), w.X3.data());
  gemm('N', 'C', naa, n, n, cplx(-1.0), w.Y1.data(), L, ZERO, w.X1.data());
  gemm('N', 'T', naa, n, n, ONE, w.Y2.data(), R, ZERO, w.X2.data());
  gemm('N', 'C', naa, n, n, cplx(-1.0), w.X3.data(), L, ONE, w.X2.data());
  gemm('N', 'H', naa, n, n, cplx(-1.0), Viaj, L, ZERO, w.X3.data());
  block_transpose(n, a, n, w.X3.data(), w.Y1.data());
  gemm('N', 'N', naa, n, n, ONE, Viaj, R, ZERO, w.X3.data());
  block_transpose(n, a, n, w.X3.data(), w.Y2.data());
  for (int nn = 0; nn < n; nn++)
    gemm('N', 'N', n2, n, a, ONE, Vija, w.X1.data() + (size_t)nn * naa, ZERO, w.Z.data() + (size_t)nn * n3);
  gemm('T', 'N', n2, a, n2, ONE, w.Z.data(), w.Y2.data(), ZERO, w.X3.data());
  for (int nn = 0; nn < n; nn++)
    gemm('N', 'N', n2, n, a, ONE, Vija, w.X2.data() + (size_t)nn * naa, ZERO, w.Z.data() + (size_t)nn * n3);
  for (size_t e = 0; e < (size_t)n2 * a; e++) w.Y1[e] += w.Y2[e];
  gemm('T', 'N', n2, a, n2, ONE, w.Z.data(), w.Y1.data(), ONE, w.X3.data());
  gemm('N', 'T', n, n, naa, cplx(-1.0), Vija, w.X3.data(), ONE, Sg.retptr(tstp, t));
}

void GF2Decomp::solve_point(int comp, int tstp, int t, const Green &Sg, const Green &G) const {
  Scratch w(nao, naux);
  if (comp == 0) { bubble_les(*this, w, tstp, t, Sg, G); exch_les(*this, w, tstp, t, Sg, G); }
  if (comp == 1) { bubble_ret(*this, w, tstp, t, Sg, G); exch_ret(*this, w, tstp, t, Sg, G); }
  if (comp == 2) { bubble_tv(*this, w, tstp, t, Sg, G); exch_tv(*this, w, tstp, t, Sg, G); }
}

// src/mol2bSolver_decomp_impl.h:412-429 (set_tstp_zero: src/greens.h:76-87).
// Per element the accumulation order (bubble, then exchange) is the
// reference's, so the threaded variant is bit-identical to the serial one.
void GF2Decomp::solve(int tstp, const Green &Sg, const Green &G, int nthreads) const {
  const int es = nao * nao, ntau = G.ntau;
  std::memset((void *)Sg.retptr(tstp, 0), 0, sizeof(cplx) * (tstp + 1) * es);
  std::memset((void *)Sg.lesptr(0, tstp), 0, sizeof(cplx) * (tstp + 1) * es);
  std::memset((void *)Sg.tvptr(tstp, 0), 0, sizeof(cplx) * (ntau + 1) * es);
  const int nles = tstp + 1, ntv = ntau + 1;
  const int total = 2 * nles + ntv;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
#endif
  {
    Scratch w(nao, naux);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int item = 0; item < total; item++) {
      if (item < nles) {
        bubble_les(*this, w, tstp, item, Sg, G);
        exch_les(*this, w, tstp, item, Sg, G);
      } else if (item < 2 * nles) {
        int t = item - nles;
        bubble_ret(*this, w, tstp, t, Sg, G);
        exch_ret(*this, w, tstp, t, Sg, G);
      } else {
        int t = item - 2 * nles;
        bubble_tv(*this, w, tstp, t, Sg, G);
        exch_tv(*this, w, tstp, t, Sg, G);
      }
    }
  }
  (void)nthreads;
}

// src/mol2bSolver_decomp_impl.h:431-670
void GF2Decomp::solve_loop(int tstp, const Green &Sg, const Green &G) const {
  const int n = nao, a = naux, es = n * n, ntau = G.ntau, sig = G.sig;
  std::memset((void *)Sg.retptr(tstp, 0), 0, sizeof(cplx) * (tstp + 1) * es);
  std::memset((void *)Sg.lesptr(0, tstp), 0, sizeof(cplx) * (tstp + 1) * es);
  std::memset((void *)Sg.tvptr(tstp, 0), 0, sizeof(cplx) * (ntau + 1) * es);
  auto V = [&](int i, int j, int q) { return Vija[((size_t)i * n + j) * a + q]; };
  // contraction kernel shared by all ten-deep loops:
  //   sum_{l n p k q m a b} f(V) * X_lk * Y_mn * W_pq
  auto contract = [&](bool exch, int i, int j, auto &&prod) {
    cplx s = 0;
    for (int l = 0; l < n; l++)
      for (int nn = 0; nn < n; nn++)
        for (int p = 0; p < n; p++)
          for (int k = 0; k < n; k++)
            for (int q = 0; q < n; q++)
              for (int m = 0; m < n; m++) {
                const cplx g = prod(l, nn, p, k, q, m);
                for (int x = 0; x < a; x++)
                  for (int y = 0; y < a; y++) {
                    double vv = exch ? V(i, l, x) * V(nn, p, x) * V(q, j, y) * V(k, m, y)
                                     : 2 * V(i, l, x) * V(nn, p, x) * V(k, j, y) * V(q, m, y);
                    s += vv * g;
                  }
              }
    return s;
  };
  for (int t = 0; t <= tstp; t++) {
    const cplx *L = G.lesptr(t, tstp), *R = G.retptr(tstp, t);
    auto les = [&](int l, int nn, int p, int k, int q, int m) {
      return L[l * n + k] * (R[m * n + nn] - std::conj(L[nn * n + m])) * L[p * n + q];
    };
    auto ret = [&](int l, int nn, int p, int k, int q, int m) {
      cplx GRlk = R[l * n + k], GLkl = std::conj(L[k * n + l]);
      cplx GRpq = R[p * n + q], GLqp = std::conj(L[q * n + p]);
      cplx GLmn = L[m * n + nn], GRnm = std::conj(R[nn * n + m]);
      return GRlk * GLmn * (GRpq - GLqp) - GLkl * (-GRnm * GLqp + GLmn * GRpq);
    };
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        Sg.lesptr(t, tstp)[i * n + j] += contract(false, i, j, les) - contract(true, i, j, les);
        Sg.retptr(tstp, t)[i * n + j] += contract(false, i, j, ret) - contract(true, i, j, ret);
      }
  }
  for (int t = 0; t <= ntau; t++) {
    const cplx *A = G.tvptr(tstp, t), *B = G.tvptr(tstp, ntau - t);
    auto tv = [&](int l, int nn, int p, int k, int q, int m) {
      return A[l * n + k] * (-(double)sig * std::conj(B[nn * n + m])) * A[p * n + q];
    };
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) Sg.tvptr(tstp, t)[i * n + j] += contract(false, i, j, tv) - contract(true, i, j, tv);
  }
}

// src/molHFSolver.cpp:81-110
void GF2Decomp::solve_HF(int tstp, cplx *hmf, const cplx *rho) const {
  const int n = nao, a = naux, n2 = n * n;
  cplx *h = hmf + (size_t)tstp * n2;
  std::vector<cplx> Xa(a), tmp((size_t)n * n2), rhoT(n2);
  gemm('N', 'N', 1, a, n2, ONE, rho, Vija, ZERO, Xa.data());
  gemm('N', 'T', 1, n2, a, cplx(2.0), Xa.data(), Vija, ONE, h);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) rhoT[i * n + j] = rho[j * n + i];
  for (int i = 0; i < n; i++) {
    gemm('N', 'T', n, n2, a, ONE, Vija + (size_t)i * n * a, Vija, ZERO, tmp.data());
    gemm('N', 'N', 1, n, n2, cplx(-1.0), rhoT.data(), tmp.data(), ONE, h + (size_t)i * n);
  }
}

// src/molHFSolver.cpp:58-79
void GF2Decomp::solve_HF_loop(int tstp, cplx *hmf, const cplx *rho) const {
  const int n = nao, a = naux;
  cplx *h = hmf + (size_t)tstp * n * n;
  auto V = [&](int i, int j, int q) { return Vija[((size_t)i * n + j) * a + q]; };
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++)
        for (int l = 0; l < n; l++)
          for (int q = 0; q < a; q++)
            h[i * n + j] += rho[k * n + l] * (2 * V(i, j, q) * V(k, l, q) - V(i, l, q) * V(k, j, q));
}

}  // namespace orc
