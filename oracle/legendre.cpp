// Gauss-Lobatto-Legendre collocation grid and the imaginary-time integrals.
// (Test infrastructure only.)
//
// The reference takes these from cpsitop (un-vendored, no pinned version;
// constructed at src/dyson.cpp:40 as Conv(nao, ntau+1), nodes read through
// Conv.collocation().x_i() at src/dyson_free.h:173, 269).  What is restated
// here is the published construction: ntau+1 GLL nodes x_i on [-1,1] with
// tau_i = beta (x_i + 1)/2, functions represented by their degree-ntau
// interpolating polynomial, and every integral evaluated exactly for those
// polynomials (Gauss-Legendre quadrature of sufficient order).  Pinned by the
// CTV1 / Ctv_tstp / Cles3_tstp / energy_conv golden values of
// tests/integrals-test.cpp:8-212.
#include "oracle.h"
#include <cmath>
#include <algorithm>

namespace orc {

static void gll_nodes(int N, std::vector<long double> &x) {
  // Newton iteration on (1-x^2) P_N'(x) with Chebyshev-Lobatto start values
  x.resize(N + 1);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int i = 0; i <= N; i++) x[i] = cosl(pi * i / N);
  std::vector<long double> P(N + 1);
  for (int it = 0; it < 100; it++) {
    long double maxd = 0;
    for (int i = 0; i <= N; i++) {
      long double xi = x[i];
      P[0] = 1; P[1] = xi;
      for (int kk = 2; kk <= N; kk++) P[kk] = ((2 * kk - 1) * xi * P[kk - 1] - (kk - 1) * P[kk - 2]) / kk;
      long double d = (xi * P[N] - P[N - 1]) / ((N + 1) * P[N]);
      x[i] = xi - d;
      maxd = std::max(maxd, fabsl(d));
    }
    if (maxd < 1e-19L) break;
  }
  std::sort(x.begin(), x.end());
  x[0] = -1; x[N] = 1;
}

static void gauss_legendre(int Q, std::vector<long double> &g, std::vector<long double> &w) {
  g.resize(Q); w.resize(Q);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int i = 0; i < Q; i++) {
    long double z = cosl(pi * (i + 0.75L) / (Q + 0.5L)), pp = 0;
    for (int it = 0; it < 100; it++) {
      long double p1 = 1, p2 = 0;
      for (int j = 1; j <= Q; j++) {
        long double p3 = p2; p2 = p1;
        p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
      }
      pp = Q * (z * p1 - p2) / (z * z - 1);
      long double z1 = z;
      z = z1 - p1 / pp;
      if (fabsl(z - z1) < 1e-19L) break;
    }
    g[i] = -z;
    w[i] = 2 / ((1 - z * z) * pp * pp);
  }
}

Legendre::Legendre(int nt) : ntau(nt), npts(nt + 1) {
  std::vector<long double> xl;
  gll_nodes(ntau, xl);
  x.resize(npts); bw.resize(npts);
  std::vector<long double> bwl(npts);
  for (int j = 0; j < npts; j++) {
    long double p = 1;
    for (int kk = 0; kk < npts; kk++) if (kk != j) p *= (xl[j] - xl[kk]);
    bwl[j] = 1 / p;
  }
  long double mx = 0;
  for (int j = 0; j < npts; j++) mx = std::max(mx, fabsl(bwl[j]));
  for (int j = 0; j < npts; j++) { x[j] = (double)xl[j]; bw[j] = (double)(bwl[j] / mx); }
  // mass matrix by Gauss-Legendre with npts points (exact to degree 2*ntau+1)
  std::vector<long double> g, w;
  gauss_legendre(npts, g, w);
  mass.assign((size_t)npts * npts, 0.0);
  std::vector<double> b(npts);
  for (int q = 0; q < npts; q++) {
    basis((double)g[q], b.data());
    for (int j = 0; j < npts; j++)
      for (int kk = 0; kk < npts; kk++) mass[(size_t)j * npts + kk] += (double)w[q] * b[j] * b[kk];
  }
}

void Legendre::basis(double xe, double *out) const {
  for (int j = 0; j < npts; j++)
    if (xe == x[j]) {
      for (int kk = 0; kk < npts; kk++) out[kk] = (kk == j) ? 1.0 : 0.0;
      return;
    }
  double s = 0;
  for (int j = 0; j < npts; j++) { out[j] = bw[j] / (xe - x[j]); s += out[j]; }
  for (int j = 0; j < npts; j++) out[j] /= s;
}

void Legendre::mixing_kernel(int nao, const cplx *B, double beta, int sig, cplx *K) const {
  const int es = nao * nao;
  std::vector<long double> g, w;
  gauss_legendre(npts, g, w);
  std::vector<double> bj(npts), bk(npts);
  std::vector<cplx> Bs(es);
  std::fill(K, K + (size_t)npts * npts * es, cplx(0));
  for (int i = 0; i < npts; i++) {
    const double ti = tau(i, beta);
    for (int piece = 0; piece < 2; piece++) {
      // piece 0: t' in [tau_i, beta], s = t' - tau_i;  piece 1: t' in [0, tau_i], s = t' - tau_i + beta
      const double a = piece == 0 ? ti : 0.0, bnd = piece == 0 ? beta : ti;
      const double len = bnd - a;
      if (len <= 0) continue;
      const double fac = piece == 0 ? 1.0 : (double)sig;
      for (int q = 0; q < npts; q++) {
        const double tp = a + len * ((double)g[q] + 1) / 2;
        const double s = piece == 0 ? tp - ti : tp - ti + beta;
        basis(2 * tp / beta - 1, bj.data());
        basis(2 * s / beta - 1, bk.data());
        std::fill(Bs.begin(), Bs.end(), cplx(0));
        for (int kk = 0; kk < npts; kk++) {
          if (bk[kk] == 0) continue;
          for (int e = 0; e < es; e++) Bs[e] += bk[kk] * B[(size_t)kk * es + e];
        }
        const double wq = fac * (double)w[q] * len / 2;
        for (int j = 0; j < npts; j++) {
          const double c = wq * bj[j];
          if (c == 0) continue;
          cplx *Kij = K + ((size_t)i * npts + j) * es;
          for (int e = 0; e < es; e++) Kij[e] += c * Bs[e];
        }
      }
    }
  }
}

}  // namespace orc
