// Gregory/BDF weights and Gauss-Lobatto-Legendre tables (host).
#include "quadrature.h"
#include <cmath>
#include <algorithm>
#include <stdexcept>
#include <thread>

namespace ned {

namespace {
typedef long double ld;

// Solve A X = B in place (n x n, nrhs columns), Gaussian elimination with
// complete pivoting in long double.
void ld_solve(int n, std::vector<ld> A, int nrhs, std::vector<ld> &B) {
  std::vector<int> colp(n);
  for (int i = 0; i < n; i++) colp[i] = i;
  for (int c = 0; c < n; c++) {
    int pr = c, pc = c;
    ld best = -1;
    for (int i = c; i < n; i++)
      for (int j = c; j < n; j++)
        if (fabsl(A[i * n + j]) > best) { best = fabsl(A[i * n + j]); pr = i; pc = j; }
    if (best <= 0) throw std::runtime_error("singular weight system");
    for (int j = 0; j < n; j++) std::swap(A[c * n + j], A[pr * n + j]);
    for (int j = 0; j < nrhs; j++) std::swap(B[c * nrhs + j], B[pr * nrhs + j]);
    for (int i = 0; i < n; i++) std::swap(A[i * n + c], A[i * n + pc]);
    std::swap(colp[c], colp[pc]);
    for (int i = c + 1; i < n; i++) {
      ld f = A[i * n + c] / A[c * n + c];
      if (f == 0) continue;
      for (int j = c; j < n; j++) A[i * n + j] -= f * A[c * n + j];
      for (int j = 0; j < nrhs; j++) B[i * nrhs + j] -= f * B[c * nrhs + j];
    }
  }
  std::vector<ld> X((size_t)n * nrhs);
  for (int i = n - 1; i >= 0; i--)
    for (int j = 0; j < nrhs; j++) {
      ld s = B[i * nrhs + j];
      for (int l = i + 1; l < n; l++) s -= A[i * n + l] * X[l * nrhs + j];
      X[i * nrhs + j] = s / A[i * n + i];
    }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < nrhs; j++) B[colp[i] * nrhs + j] = X[i * nrhs + j];
}

// Lagrange-basis coefficients on the integer nodes 0..k: P[a][l] is the
// coefficient of x^a in the basis polynomial of node l (inverse Vandermonde).
std::vector<ld> lagrange_coeffs(int k) {
  int k1 = k + 1;
  std::vector<ld> V((size_t)k1 * k1), P((size_t)k1 * k1, 0);
  for (int i = 0; i <= k; i++) {
    ld p = 1;
    for (int a = 0; a <= k; a++) { V[i * k1 + a] = p; p *= i; }
  }
  for (int i = 0; i < k1; i++) P[i * k1 + i] = 1;
  ld_solve(k1, V, k1, P);
  return P;
}
}  // namespace

Weights::Weights(int kk) : k(kk) {
  if (k < 1) throw std::invalid_argument("integrator order k must be >= 1");
  const int k1 = k + 1;
  std::vector<ld> P = lagrange_coeffs(k);
  interp.resize(k1 * k1);
  for (int i = 0; i < k1 * k1; i++) interp[i] = (double)P[i];
  // derivative of basis l at node n
  auto dfun = [](const std::vector<ld> &Pc, int kk1, int nn, int l) {
    ld s = 0, pw = 1;
    for (int a = 1; a < kk1; a++) { s += Pc[a * kk1 + l] * a * pw; pw *= nn; }
    return s;
  };
  diff.resize(k1 * k1);
  for (int nn = 0; nn <= k; nn++)
    for (int l = 0; l <= k; l++) diff[nn * k1 + l] = (double)dfun(P, k1, nn, l);
  // integral of basis l from m to n
  integ.resize(k1 * k1 * k1);
  for (int m = 0; m <= k; m++)
    for (int nn = 0; nn <= k; nn++)
      for (int l = 0; l <= k; l++) {
        ld s = 0, pm = m, pn = nn;
        for (int a = 0; a <= k; a++) { s += P[a * k1 + l] * (pn - pm) / (a + 1); pm *= m; pn *= nn; }
        integ[(m * k1 + nn) * k1 + l] = (double)s;
      }
  // backward differentiation: minus the derivative at node 0 of the order-k+1 basis
  {
    std::vector<ld> P2 = lagrange_coeffs(k + 1);
    bd.resize(k + 2);
    for (int l = 0; l < k + 2; l++) bd[l] = (double)(-dfun(P2, k + 2, 0, l));
  }
  // start weights: integral of basis l from 0 to n
  start.resize(k1 * k1);
  for (int nn = 0; nn <= k; nn++)
    for (int l = 0; l <= k; l++) {
      ld s = 0, pn = nn;
      for (int a = 0; a <= k; a++) { s += P[a * k1 + l] * pn / (a + 1); pn *= nn; }
      start[nn * k1 + l] = (double)s;
    }
  // Gregory end corrections (src/integration.cpp:143-171)
  {
    std::vector<ld> A((size_t)k1 * k1, 0), y(k1);
    for (int i = 0; i <= k; i++) {
      y[i] = ((i + 1) % 2 ? -1.0L : 1.0L) / (i + 2);
      for (int j = 0; j <= i; j++) A[i * k1 + j] = ((i + j) % 2 ? -1.0L : 1.0L) / (i - j + 1);
    }
    ld_solve(k1, A, 1, y);
    Omega.resize(k1 * k1);
    for (int nn = 0; nn <= k; nn++) {
      std::vector<ld> row(k + nn + 2, 1.0L);
      for (int i = 0; i <= k; i++) {
        ld binom = 1;
        for (int j = 0; j <= i; j++) {
          if (j > 0) binom = binom * (i - j + 1) / j;
          ld x = ((i + j) % 2 ? -1.0L : 1.0L) * y[i] * binom;
          row[j] += x;
          row[k + nn - j + 1] += x;
        }
      }
      for (int i = 0; i <= k; i++) Omega[nn * k1 + i] = (double)row[i];
    }
  }
  // extrapolation to node -1 (src/integration.cpp:162-170, from the double table)
  ex.assign(k1, 0.0);
  for (int l = 0; l <= k; l++)
    for (int j = 0; j <= k; j++) ex[l] += interp[j * k1 + l] * ((j % 2) ? -1.0 : 1.0);
}

double Weights::gregory(int n, int j) const {
  const int k1 = k + 1;
  if (n <= k && j <= k) return start[n * k1 + j];
  if (n <= 2 * k + 1) return j <= k ? Omega[(n - k - 1) * k1 + j] : omega(n - j);
  if (j <= k) return omega(j);
  if (j < n - k) return 1.0;
  return omega(n - j);
}

namespace {
void gauss_legendre(int q, std::vector<ld> &g, std::vector<ld> &w) {
  const ld pi = 3.141592653589793238462643383279502884L;
  g.resize(q); w.resize(q);
  for (int i = 0; i < q; i++) {
    ld z = cosl(pi * (i + 0.75L) / (q + 0.5L)), dp = 1;
    for (int it = 0; it < 100; it++) {
      ld p0 = 1, p1 = z;
      for (int l = 2; l <= q; l++) { ld p2 = ((2 * l - 1) * z * p1 - (l - 1) * p0) / l; p0 = p1; p1 = p2; }
      if (q == 1) { p1 = z; p0 = 1; }
      dp = q * (z * p1 - p0) / (z * z - 1);
      ld dz = p1 / dp;
      z -= dz;
      if (fabsl(dz) < 1e-20L) break;
    }
    g[i] = -z;
    w[i] = 2 / ((1 - z * z) * dp * dp);
  }
}
}  // namespace

GLL::GLL(int nt) : ntau(nt), np(nt + 1) {
  if (ntau < 1) throw std::invalid_argument("ntau must be >= 1");
  const ld pi = 3.141592653589793238462643383279502884L;
  std::vector<ld> xl(np);
  // interior nodes: roots of P_N'; Newton on (1-x^2) P_N'(x) from Chebyshev-Lobatto points
  for (int i = 0; i < np; i++) {
    ld z = -cosl(pi * i / ntau);
    for (int it = 0; it < 100; it++) {
      ld p0 = 1, p1 = z;
      for (int l = 2; l <= ntau; l++) { ld p2 = ((2 * l - 1) * z * p1 - (l - 1) * p0) / l; p0 = p1; p1 = p2; }
      ld dz = (z * p1 - p0) / (np * p1);
      z -= dz;
      if (fabsl(dz) < 1e-20L) break;
    }
    xl[i] = z;
  }
  xl[0] = -1; xl[ntau] = 1;
  std::sort(xl.begin(), xl.end());
  x.resize(np); bw.resize(np);
  std::vector<ld> wl(np);
  ld mx = 0;
  for (int j = 0; j < np; j++) {
    ld p = 1;
    for (int m = 0; m < np; m++) if (m != j) p *= (xl[j] - xl[m]);
    wl[j] = 1 / p;
    mx = std::max(mx, fabsl(wl[j]));
  }
  for (int j = 0; j < np; j++) { x[j] = (double)xl[j]; bw[j] = (double)(wl[j] / mx); }
  std::vector<ld> g, w;
  gauss_legendre(np, g, w);
  mass.assign((size_t)np * np, 0.0);
  std::vector<double> b(np);
  for (int q = 0; q < np; q++) {
    basis((double)g[q], b.data());
    for (int j = 0; j < np; j++)
      for (int m = 0; m < np; m++) mass[(size_t)j * np + m] += (double)w[q] * b[j] * b[m];
  }
}

void GLL::basis(double xe, double *out) const {
  for (int j = 0; j < np; j++)
    if (xe == x[j]) {
      for (int m = 0; m < np; m++) out[m] = (m == j);
      return;
    }
  double s = 0;
  for (int j = 0; j < np; j++) { out[j] = bw[j] / (xe - x[j]); s += out[j]; }
  for (int j = 0; j < np; j++) out[j] /= s;
}

void GLL::mixing_rows(double beta, int sig, std::vector<double> &la, std::vector<double> &lb) const {
  const int nq = 2 * np;
  std::vector<ld> g, w;
  gauss_legendre(np, g, w);
  la.assign((size_t)np * nq * np, 0.0);
  lb.assign((size_t)np * nq * np, 0.0);
  std::vector<double> bj(np), bk(np);
  for (int i = 0; i < np; i++) {
    const double ti = beta * (x[i] + 1) / 2;
    for (int piece = 0; piece < 2; piece++) {
      const double lo = piece ? 0.0 : ti, hi = piece ? ti : beta, len = hi - lo;
      const double fac = piece ? (double)sig : 1.0;
      for (int q = 0; q < np; q++) {
        const int qq = piece * np + q;
        double *ra = &la[((size_t)i * nq + qq) * np], *rb = &lb[((size_t)i * nq + qq) * np];
        if (len <= 0) continue;  // zero-length piece at the end points
        const double t = lo + len * ((double)g[q] + 1) / 2;
        const double s = piece ? t - ti + beta : t - ti;
        basis(2 * t / beta - 1, bj.data());
        basis(2 * s / beta - 1, bk.data());
        const double wq = fac * (double)w[q] * len / 2;
        for (int j = 0; j < np; j++) { ra[j] = wq * bj[j]; rb[j] = bk[j]; }
      }
    }
  }
}

void GLL::conv_weights(double beta, int sig, std::vector<double> &w) const {
  std::vector<ld> g, gw;
  gauss_legendre(np, g, gw);
  w.assign((size_t)np * np * np, 0.0);
  auto rows = [&](int m0, int m1) {
    std::vector<double> ba((size_t)np * np), bb((size_t)np * np), wq(np);
    for (int m = m0; m < m1; m++) {
      const double tm = beta * (x[m] + 1) / 2;
      double *wm = &w[(size_t)m * np * np];
      for (int piece = 0; piece < 2; piece++) {
        // piece 0: t in [0, tau_m], A at tau_m - t; piece 1: t in [tau_m, beta], A at tau_m - t + beta
        const double lo = piece ? tm : 0.0, hi = piece ? beta : tm, len = hi - lo;
        if (len <= 0) continue;
        const double fac = piece ? (double)sig : 1.0;
        for (int q = 0; q < np; q++) {
          const double t = lo + len * ((double)g[q] + 1) / 2;
          const double s = piece ? tm - t + beta : tm - t;
          basis(2 * s / beta - 1, &ba[(size_t)q * np]);
          basis(2 * t / beta - 1, &bb[(size_t)q * np]);
          wq[q] = fac * (double)gw[q] * len / 2;
        }
        for (int q = 0; q < np; q++)
          for (int j = 0; j < np; j++) {
            const double a = wq[q] * ba[(size_t)q * np + j];
            const double *b = &bb[(size_t)q * np];
            double *o = wm + (size_t)j * np;
            for (int k = 0; k < np; k++) o[k] += a * b[k];
          }
      }
    }
  };
  const int nth = std::max(1, std::min<int>(16, (int)std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  const int per = (np + nth - 1) / nth;
  for (int t = 0; t < nth; t++) {
    const int m0 = t * per, m1 = std::min(np, m0 + per);
    if (m0 < m1) th.emplace_back(rows, m0, m1);
  }
  for (auto &t : th) t.join();
}

}  // namespace ned

namespace ned {
void sym_eig(int n, const double *Ain, double *w, double *V) {
  std::vector<double> A(Ain, Ain + (size_t)n * n);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) V[i * n + j] = (i == j);
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0, tot = 0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        tot += A[i * n + j] * A[i * n + j];
        if (i != j) off += A[i * n + j] * A[i * n + j];
      }
    if (off <= 1e-32 * tot || off == 0) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = A[p * n + q];
        if (apq == 0) continue;
        const double theta = (A[q * n + q] - A[p * n + p]) / (2 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        const double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; k++) {
          double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          double vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = c * vkp - s * vkq;
          V[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; i++) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return A[a * n + a] < A[b * n + b]; });
  std::vector<double> Vs((size_t)n * n);
  for (int c = 0; c < n; c++) {
    w[c] = A[idx[c] * n + idx[c]];
    for (int r = 0; r < n; r++) Vs[r * n + c] = V[r * n + idx[c]];
  }
  std::copy(Vs.begin(), Vs.end(), V);
}
}  // namespace ned
