// Gregory / BDF / polynomial weights, restating src/integration.cpp:8-171 and
// the regime switch of src/integration.h:78-91.  Built in long double and
// rounded to double, as the reference does.  (Test infrastructure only.)
#include "oracle.h"
#include <cmath>

namespace orc {

// src/integration.cpp:8-27: P = inverse of the Vandermonde matrix M(i,n) = i^n
static void poly_interp_ld(int k, std::vector<long double> &P) {
  int k1 = k + 1;
  std::vector<long double> M((size_t)k1 * k1), Inv((size_t)k1 * k1, 0.0L);
  for (int i = 0; i <= k; i++) {
    M[i * k1 + 0] = 1.0L;
    for (int n = 1; n <= k; n++) M[i * k1 + n] = M[i * k1 + n - 1] * i;
  }
  for (int i = 0; i < k1; i++) Inv[i * k1 + i] = 1.0L;
  fullpiv_solve_ld(k1, M.data(), k1, Inv.data());
  P = Inv;
}

// src/integration.cpp:38-53: derivative of basis l at node n
static void poly_diff_ld(int k, const std::vector<long double> &P, std::vector<long double> &D) {
  int k1 = k + 1;
  D.assign((size_t)k1 * k1, 0.0L);
  for (int n = 0; n <= k; n++)
    for (int l = 0; l <= k; l++) {
      long double npow = 1.0L, x = 0.0L;
      for (int a = 1; a <= k; a++) {
        if (a > 1) npow *= n;
        x += P[a * k1 + l] * a * npow;
      }
      D[n * k1 + l] = x;
    }
}

// src/integration.cpp:67-86: integral of basis l from m to n
static void poly_integ_ld(int k, const std::vector<long double> &P, std::vector<long double> &Iv) {
  int k1 = k + 1;
  Iv.assign((size_t)k1 * k1 * k1, 0.0L);
  for (int m = 0; m <= k; m++)
    for (int n = 0; n <= k; n++)
      for (int l = 0; l <= k; l++) {
        long double mpow = m, npow = n, x = 0.0L;
        for (int a = 0; a <= k; a++) {
          x += P[a * k1 + l] * (npow - mpow) / (a + 1);
          mpow *= m;
          npow *= n;
        }
        Iv[(m * k1 + n) * k1 + l] = x;
      }
}

static long double binom(int row, int col) {
  long double r = 1.0L;
  for (int i = 1; i <= col; i++) r = r * (row - col + i) / i;
  return r;
}

Integrator::Integrator(int kk) : k(kk) {
  int k1 = k + 1;
  std::vector<long double> P, D, Iv;
  poly_interp_ld(k, P);
  poly_diff_ld(k, P, D);
  poly_integ_ld(k, P, Iv);
  poly_interp_.resize(k1 * k1);
  poly_diff_.resize(k1 * k1);
  poly_integ_.resize(k1 * k1 * k1);
  for (int i = 0; i < k1 * k1; i++) { poly_interp_[i] = (double)P[i]; poly_diff_[i] = (double)D[i]; }
  for (int i = 0; i < k1 * k1 * k1; i++) poly_integ_[i] = (double)Iv[i];

  // src/integration.cpp:100-112: BD[l] = -D_{k+1}(0, l)
  {
    std::vector<long double> P2, D2;
    poly_interp_ld(k + 1, P2);
    poly_diff_ld(k + 1, P2, D2);
    bd_.resize(k + 2);
    for (int i = 0; i < k + 2; i++) bd_[i] = (double)(-D2[i]);
  }
  // src/integration.cpp:114-132: start weights s[n][l] = int_0^n basis_l
  start_.resize(k1 * k1);
  for (int n = 0; n <= k; n++)
    for (int l = 0; l <= k; l++) {
      long double x = 0.0L, npow = n;
      for (int a = 0; a <= k; a++) {
        x += P[a * k1 + l] * npow / (a + 1.0L);
        npow *= n;
      }
      start_[n * k1 + l] = (double)x;
    }
  // src/integration.cpp:143-171: Gregory end corrections Omega (last row = omega)
  {
    std::vector<long double> A((size_t)k1 * k1, 0.0L), y(k1);
    for (int i = 0; i <= k; i++) {
      y[i] = (long double)(1 - 2 * ((i + 1) % 2)) / (i + 2.0L);
      for (int j = 0; j <= i; j++) A[i * k1 + j] = (long double)(1 - 2 * ((i + j) % 2)) / (i - j + 1.0L);
    }
    fullpiv_solve_ld(k1, A.data(), 1, y.data());
    Omega_.resize(k1 * k1);
    for (int n = 0; n <= k; n++) {
      std::vector<long double> Iv2(k + n + 2, 1.0L);
      for (int i = 0; i <= k; i++)
        for (int j = 0; j <= i; j++) {
          long double x = (long double)(1 - 2 * ((i + j) % 2)) * y[i] * binom(i, j);
          Iv2[j] += x;
          Iv2[k + n - j + 1] += x;
        }
      for (int i = 0; i <= k; i++) Omega_[n * k1 + i] = (double)Iv2[i];
    }
  }
  // src/integration.cpp:162-170 (make_ex_weights, from the double poly_interp)
  ex_.assign(k1, 0.0);
  for (int l = 0; l <= k; l++)
    for (int j = 0; j <= k; j++) ex_[l] += poly_interp_[j * k1 + l] * (1 - 2 * (j % 2));
}

// src/integration.h:78-91
double Integrator::gregory_weights(int n, int j) const {
  int k1 = k + 1;
  if (n <= k && j <= k) return start_[n * k1 + j];
  if (n <= 2 * k + 1) {
    if (j <= k) return Omega_[(n - k - 1) * k1 + j];
    return omega(n - j);
  }
  if (j <= k) return omega(j);
  if (j < n - k) return 1.0;
  return omega(n - j);
}

}  // namespace orc
