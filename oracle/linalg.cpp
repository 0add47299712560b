// Dense helpers for the CPU oracle (test infrastructure only): they stand in
// for the Eigen expressions and Eigen::FullPivLU calls of the reference.
#include "oracle.h"
#include <cmath>
#include <algorithm>
#include <cstring>
#include <stdexcept>

namespace orc {

static inline double cj(double x) { return x; }
static inline cplx cj(const cplx &x) { return std::conj(x); }

template <class T>
static void materialize(char op, int R, int Cc, const T *S, std::vector<T> &out) {
  // out = op(S) as an R x Cc row-major matrix
  out.resize((size_t)R * Cc);
  switch (op) {
    case 'N':
      std::memcpy(out.data(), S, sizeof(T) * (size_t)R * Cc);
      break;
    case 'C':
      for (size_t i = 0; i < (size_t)R * Cc; i++) out[i] = cj(S[i]);
      break;
    case 'T':
      for (int r = 0; r < R; r++)
        for (int c = 0; c < Cc; c++) out[(size_t)r * Cc + c] = S[(size_t)c * R + r];
      break;
    case 'H':
      for (int r = 0; r < R; r++)
        for (int c = 0; c < Cc; c++) out[(size_t)r * Cc + c] = cj(S[(size_t)c * R + r]);
      break;
    default:
      throw std::runtime_error("gemm: bad op");
  }
}

template <class TA, class TB>
void gemm(char opA, char opB, int M, int N, int K, cplx alpha, const TA *A, const TB *B,
          cplx beta, cplx *C) {
  std::vector<TA> a;
  std::vector<TB> b;
  materialize(opA, M, K, A, a);
  materialize(opB, K, N, B, b);
  if (beta == cplx(0.0)) {
    std::fill(C, C + (size_t)M * N, cplx(0.0));
  } else if (beta != cplx(1.0)) {
    for (size_t i = 0; i < (size_t)M * N; i++) C[i] *= beta;
  }
  for (int i = 0; i < M; i++) {
    cplx *c = C + (size_t)i * N;
    for (int kk = 0; kk < K; kk++) {
      const cplx aik = alpha * cplx(a[(size_t)i * K + kk]);
      const TB *brow = b.data() + (size_t)kk * N;
      for (int j = 0; j < N; j++) c[j] += aik * brow[j];
    }
  }
}

template void gemm<cplx, cplx>(char, char, int, int, int, cplx, const cplx *, const cplx *, cplx, cplx *);
template void gemm<double, cplx>(char, char, int, int, int, cplx, const double *, const cplx *, cplx, cplx *);
template void gemm<cplx, double>(char, char, int, int, int, cplx, const cplx *, const double *, cplx, cplx *);
template void gemm<double, double>(char, char, int, int, int, cplx, const double *, const double *, cplx, cplx *);

template <class T, class R>
static void fullpiv_solve_t(int n, const T *Min, int nrhs, T *Q) {
  std::vector<T> A(Min, Min + (size_t)n * n);
  std::vector<int> rowp(n), colp(n);
  for (int i = 0; i < n; i++) colp[i] = i;
  // LU with complete pivoting: P A Qc = L U
  std::vector<int> perm_r(n);
  for (int i = 0; i < n; i++) perm_r[i] = i;
  for (int kk = 0; kk < n; kk++) {
    R best = -1;
    int br = kk, bc = kk;
    for (int i = kk; i < n; i++)
      for (int j = kk; j < n; j++) {
        R v = std::abs(A[(size_t)i * n + j]);
        if (v > best) { best = v; br = i; bc = j; }
      }
    if (br != kk) {
      for (int j = 0; j < n; j++) std::swap(A[(size_t)kk * n + j], A[(size_t)br * n + j]);
      std::swap(perm_r[kk], perm_r[br]);
    }
    if (bc != kk) {
      for (int i = 0; i < n; i++) std::swap(A[(size_t)i * n + kk], A[(size_t)i * n + bc]);
      std::swap(colp[kk], colp[bc]);
    }
    T piv = A[(size_t)kk * n + kk];
    if (piv == T(0)) continue;  // singular block: leave as is (Eigen returns a least-effort solution)
    for (int i = kk + 1; i < n; i++) {
      T f = A[(size_t)i * n + kk] / piv;
      A[(size_t)i * n + kk] = f;
      if (f == T(0)) continue;
      for (int j = kk + 1; j < n; j++) A[(size_t)i * n + j] -= f * A[(size_t)kk * n + j];
    }
  }
  // apply row permutation to Q
  std::vector<T> B((size_t)n * nrhs), Y((size_t)n * nrhs);
  for (int i = 0; i < n; i++)
    for (int c = 0; c < nrhs; c++) B[(size_t)i * nrhs + c] = Q[(size_t)perm_r[i] * nrhs + c];
  // forward (unit L)
  for (int i = 0; i < n; i++)
    for (int j = 0; j < i; j++) {
      T f = A[(size_t)i * n + j];
      if (f == T(0)) continue;
      for (int c = 0; c < nrhs; c++) B[(size_t)i * nrhs + c] -= f * B[(size_t)j * nrhs + c];
    }
  // backward (U)
  for (int i = n - 1; i >= 0; i--) {
    for (int j = i + 1; j < n; j++) {
      T f = A[(size_t)i * n + j];
      if (f == T(0)) continue;
      for (int c = 0; c < nrhs; c++) B[(size_t)i * nrhs + c] -= f * B[(size_t)j * nrhs + c];
    }
    T d = A[(size_t)i * n + i];
    for (int c = 0; c < nrhs; c++) B[(size_t)i * nrhs + c] = (d == T(0)) ? T(0) : B[(size_t)i * nrhs + c] / d;
  }
  // undo column permutation: x[colp[i]] = y[i]
  for (int i = 0; i < n; i++)
    for (int c = 0; c < nrhs; c++) Q[(size_t)colp[i] * nrhs + c] = B[(size_t)i * nrhs + c];
}

void fullpiv_solve(int n, const cplx *M, int nrhs, cplx *Q) { fullpiv_solve_t<cplx, double>(n, M, nrhs, Q); }
void fullpiv_solve_ld(int n, const long double *M, int nrhs, long double *Q) {
  fullpiv_solve_t<long double, long double>(n, M, nrhs, Q);
}

void symeig(int n, const double *Ain, double *evals, double *V) {
  std::vector<double> A(Ain, Ain + (size_t)n * n);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) V[(size_t)i * n + j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0;
    for (int i = 0; i < n; i++)
      for (int j = i + 1; j < n; j++) off += A[(size_t)i * n + j] * A[(size_t)i * n + j];
    if (off < 1e-30) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        double apq = A[(size_t)p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
        double theta = (aqq - app) / (2 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int kk = 0; kk < n; kk++) {
          double akp = A[(size_t)kk * n + p], akq = A[(size_t)kk * n + q];
          A[(size_t)kk * n + p] = c * akp - s * akq;
          A[(size_t)kk * n + q] = s * akp + c * akq;
        }
        for (int kk = 0; kk < n; kk++) {
          double apk = A[(size_t)p * n + kk], aqk = A[(size_t)q * n + kk];
          A[(size_t)p * n + kk] = c * apk - s * aqk;
          A[(size_t)q * n + kk] = s * apk + c * aqk;
        }
        for (int kk = 0; kk < n; kk++) {
          double vkp = V[(size_t)kk * n + p], vkq = V[(size_t)kk * n + q];
          V[(size_t)kk * n + p] = c * vkp - s * vkq;
          V[(size_t)kk * n + q] = s * vkp + c * vkq;
        }
      }
  }
  // sort ascending
  std::vector<int> idx(n);
  for (int i = 0; i < n; i++) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return A[(size_t)a * n + a] < A[(size_t)b * n + b]; });
  std::vector<double> Vs((size_t)n * n);
  for (int c = 0; c < n; c++) {
    evals[c] = A[(size_t)idx[c] * n + idx[c]];
    for (int r = 0; r < n; r++) Vs[(size_t)r * n + c] = V[(size_t)r * n + idx[c]];
  }
  std::memcpy(V, Vs.data(), sizeof(double) * n * n);
}

GreenStore::GreenStore(int nt, int ntau, int nao, int sig) {
  size_t es = (size_t)nao * nao;
  size_t ntri = (size_t)(nt + 1) * (nt + 2) / 2;
  ret.assign(ntri * es, cplx(0));
  les.assign(ntri * es, cplx(0));
  tv.assign((size_t)(nt + 1) * (ntau + 1) * es, cplx(0));
  mat.assign((size_t)(ntau + 1) * es, cplx(0));
  g.nt = nt; g.ntau = ntau; g.nao = nao; g.sig = sig;
  g.ret = ret.data(); g.les = les.data(); g.tv = tv.data(); g.mat = mat.data();
}

}  // namespace orc
