// Host-side time and imaginary-time quadrature tables.
#pragma once
#include <vector>

namespace ned {

// Integration::Integrator(k) (src/integration.h:23-124,
// src/integration.cpp:8-171), built in long double and rounded to double.
struct Weights {
  int k = 0;
  std::vector<double> interp, diff, integ, bd, start, Omega, ex;
  explicit Weights(int k);
  double omega(int j) const { return Omega[k * (k + 1) + j]; }
  double gregory(int n, int j) const;  // src/integration.h:78-91
};

// Gauss-Lobatto-Legendre collocation with ntau+1 nodes on [-1,1].
struct GLL {
  int ntau = 0, np = 0;
  std::vector<double> x, bw;   // nodes (ascending), barycentric weights
  std::vector<double> mass;    // np x np: int_{-1}^{1} l_j l_k dx
  explicit GLL(int ntau);
  void basis(double xe, double *out) const;
  // Quadrature description of the antiperiodic mixing integral
  //   C(tau_i) = int_0^beta A(t) B(t - tau_i) dt,  B(s<0) = sig B(s+beta):
  // for every node i, nq = 2*np points t_q with weight w_q (sig folded in),
  // interpolation rows la[i][q][j] = w_q l_j(t_q) and lb[i][q][k] = l_k(s_q).
  void mixing_rows(double beta, int sig, std::vector<double> &la, std::vector<double> &lb) const;
  // Weights of the antiperiodic imaginary-time convolution
  //   C(tau_m) = int_0^beta A(tau_m - t) B(t) dt,  A(s<0) = sig A(s+beta),
  // of the nodal interpolants of A and B:
  //   C_m = sum_{j,k} w[(m*np + j)*np + k] A_j B_k.
  // The integrand is a polynomial of degree 2*ntau on [0, tau_m] and on
  // [tau_m, beta], so np-point Gauss-Legendre on each piece is exact.
  void conv_weights(double beta, int sig, std::vector<double> &w) const;
};

}  // namespace ned

namespace ned {
// Real symmetric eigendecomposition (cyclic Jacobi): A = V diag(w) V^T,
// eigenvalues ascending, eigenvectors in the columns of V (row-major n x n).
void sym_eig(int n, const double *A, double *w, double *V);
}  // namespace ned
