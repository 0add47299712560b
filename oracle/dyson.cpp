// Dyson timestep, bootstrap, history integrals, predictor, energy and free
// Green's function, restating src/dyson*.h and src/dyson.cpp.
// (Test infrastructure only.)
#include "oracle.h"
#include <cmath>
#include <cstring>
#include <algorithm>

namespace orc {

namespace {
const cplx ONE(1.0), ZERO(0.0), I1(0.0, 1.0);
double fro(const cplx *a, const cplx *b, int es) {
  double s = 0;
  for (int e = 0; e < es; e++) s += std::norm(a[e] - b[e]);
  return std::sqrt(s);
}
}  // namespace

Dyson::Dyson(int nt_, int ntau_, int nao_, int k_)
    : nt(nt_), ntau(ntau_), nao(nao_), es(nao_ * nao_), k(k_), I(k_), L(ntau_),
      Q((size_t)(nt_ + 1) * es), X((size_t)(nt_ + 1) * es), M((size_t)k_ * k_ * es),
      NTauTmp((size_t)(k_ + 1) * (ntau_ + 1) * es), tmp(es) {}

const cplx *Dyson::mixing_kernel(const Green &B, double beta) const {
  const size_t nb = (size_t)(ntau + 1) * es;
  if (mixK.empty() || mixBeta != beta || mixSig != B.sig || mixB.size() != nb ||
      std::memcmp(mixB.data(), B.mat, sizeof(cplx) * nb) != 0) {
    mixB.assign(B.mat, B.mat + nb);
    mixBeta = beta;
    mixSig = B.sig;
    mixK.assign((size_t)(ntau + 1) * (ntau + 1) * es, cplx(0));
    L.mixing_kernel(nao, B.mat, beta, B.sig, mixK.data());
  }
  return mixK.data();
}

void Dyson::conv_mixing(cplx *C, const cplx *A, const Green &B, double beta) const {
  const cplx *K = mixing_kernel(B, beta);
  const int np = ntau + 1;
  std::fill(C, C + (size_t)np * es, cplx(0));
  for (int i = 0; i < np; i++)
    for (int j = 0; j < np; j++)
      gemm('N', 'N', nao, nao, nao, ONE, A + (size_t)j * es, K + ((size_t)i * np + j) * es, ONE, C + (size_t)i * es);
}

void Dyson::conv_lesser(cplx *res, const cplx *A, const cplx *B, double beta) const {
  const int np = ntau + 1;
  std::vector<cplx> Bw((size_t)np * es, cplx(0));
  for (int j = 0; j < np; j++)
    for (int kk = 0; kk < np; kk++) {
      double w = 0.5 * beta * L.mass[(size_t)j * np + kk];
      for (int e = 0; e < es; e++) Bw[(size_t)j * es + e] += w * B[(size_t)kk * es + e];
    }
  for (int j = 0; j < np; j++) gemm('N', 'N', nao, nao, nao, ONE, A + (size_t)j * es, Bw.data() + (size_t)j * es, ONE, res);
}

cplx Dyson::conv_energy(const cplx *S, const cplx *G, double beta, int sig) const {
  const int np = ntau + 1;
  cplx r = 0;
  for (int j = 0; j < np; j++)
    for (int kk = 0; kk < np; kk++) {
      double w = 0.5 * beta * L.mass[(size_t)j * np + (ntau - kk)];
      if (w == 0) continue;
      cplx s = 0;
      for (int e = 0; e < es; e++) s += S[(size_t)j * es + e] * std::conj(G[(size_t)kk * es + e]);
      r += w * s;
    }
  return -(double)sig * r;
}

// src/dyson.cpp:47-117
void Dyson::Extrapolate(int n, const Green &G) const {
  std::vector<double> exw(k + 1, 0.0);
  for (int l = 0; l <= k; l++)
    for (int j = 0; j <= k; j++) exw[l] += I.poly_interp(j, l) * (1 - 2 * (j % 2));
  std::memset((void *)G.tvptr(n, 0), 0, sizeof(cplx) * es * (ntau + 1));
  for (int l = 0; l <= ntau; l++) {
    cplx *r = G.tvptr(n, l);
    for (int j = 0; j <= k; j++) {
      const cplx *s = G.tvptr(n - j - 1, l);
      for (int e = 0; e < es; e++) r[e] += exw[j] * s[e];
    }
  }
  std::memset((void *)G.retptr(n, 0), 0, sizeof(cplx) * (n + 1) * es);
  for (int l = 0; l < n - k; l++) {
    cplx *r = G.retptr(n, n - l);
    for (int j = 0; j <= k; j++) {
      const cplx *s = G.retptr(n - j - 1, n - l - j - 1);
      for (int e = 0; e < es; e++) r[e] += exw[j] * s[e];
    }
  }
  for (int l = 0; l <= k; l++) {
    int jcut = (l <= n - k - 1) ? k : (n - l - 1);
    cplx *r = G.retptr(n, l);
    for (int j = 0; j <= jcut; j++) {
      const cplx *s = G.retptr(n - j - 1, l);
      for (int e = 0; e < es; e++) r[e] += exw[j] * s[e];
    }
    for (int j = jcut + 1; j <= k; j++) {
      const cplx *s = G.retptr(l, n - j - 1);
      for (int a = 0; a < nao; a++)
        for (int b = 0; b < nao; b++) r[a * nao + b] -= exw[j] * std::conj(s[b * nao + a]);
    }
  }
  std::memset((void *)G.lesptr(0, n), 0, sizeof(cplx) * (n + 1) * es);
  {
    cplx *r = G.lesptr(0, n);
    const cplx *s = G.tvptr(n, 0);
    for (int a = 0; a < nao; a++)
      for (int b = 0; b < nao; b++) r[a * nao + b] = -std::conj(s[b * nao + a]);
  }
  for (int l = 1; l <= k; l++) {
    int jcut = (k > n - l - 1) ? (n - l - 1) : k;
    cplx *r = G.lesptr(l, n);
    for (int j = 0; j <= jcut; j++) {
      const cplx *s = G.lesptr(l, n - 1 - j);
      for (int e = 0; e < es; e++) r[e] += exw[j] * s[e];
    }
    for (int j = jcut + 1; j <= k; j++) {
      const cplx *s = G.lesptr(n - 1 - j, l);
      for (int a = 0; a < nao; a++)
        for (int b = 0; b < nao; b++) r[a * nao + b] -= exw[j] * std::conj(s[b * nao + a]);
    }
  }
  for (int l = k + 1; l <= n; l++) {
    cplx *r = G.lesptr(l, n);
    for (int j = 0; j <= k; j++) {
      const cplx *s = G.lesptr(l - j - 1, n - j - 1);
      for (int e = 0; e < es; e++) r[e] += exw[j] * s[e];
    }
  }
}

// src/dyson_integrals.h:80-116
void Dyson::CTV1(cplx *ctv, const Green &A, const Green &Acc, const Green &B, int n, double dt) const {
  int ntop = (n > k) ? n : k;
  for (int j = 0; j <= ntop; j++) {
    if (j > n) {
      const cplx *s = Acc.retptr(j, n);
      double w = -dt * I.gregory_weights(n, j);
      for (int a = 0; a < nao; a++)
        for (int b = 0; b < nao; b++) tmp[a * nao + b] = w * std::conj(s[b * nao + a]);
      for (int m = 0; m <= ntau; m++)
        gemm('N', 'N', nao, nao, nao, ONE, tmp.data(), B.tvptr(j, m), ONE, ctv + (size_t)m * es);
    } else {
      const cplx *Am = A.retptr(n, j);
      cplx w = dt * I.gregory_weights(n, j);
      for (int m = 0; m <= ntau; m++) gemm('N', 'N', nao, nao, nao, w, Am, B.tvptr(j, m), ONE, ctv + (size_t)m * es);
    }
  }
}

// src/dyson_integrals.h:254-324
void Dyson::Ctv_tstp(int tstp, const Green &C, const Green &A, const Green &Acc, const Green &B, const Green &Bcc,
                     double beta, double dt) const {
  (void)Bcc;
  conv_mixing(NTauTmp.data(), A.tvptr(tstp, 0), B, beta);
  CTV1(NTauTmp.data(), A, Acc, B, tstp, dt);
  std::memcpy((void *)C.tvptr(tstp, 0), NTauTmp.data(), sizeof(cplx) * (ntau + 1) * es);
}

// src/dyson_integrals.h:336-380
void Dyson::Cles2_tstp(int j1, int j2, const Green &A, const Green &Acc, const Green &B, const Green &Bcc, int m,
                       double dt, cplx *res) const {
  int top = m > k ? m : k;
  for (int j = 0; j <= top; j++) {
    cplx *x = X.data() + (size_t)j * es;
    if (m >= j) {
      const cplx *s = Bcc.retptr(m, j);
      double w = dt * I.gregory_weights(m, j);
      for (int a = 0; a < nao; a++)
        for (int b = 0; b < nao; b++) x[a * nao + b] = w * std::conj(s[b * nao + a]);
    } else {
      const cplx *s = B.retptr(j, m);
      double w = -dt * I.gregory_weights(m, j);
      for (int e = 0; e < es; e++) x[e] = w * s[e];
    }
  }
  for (int n = j1; n <= j2; n++) {
    cplx *r = res + (size_t)(n - j1) * es;
    for (int j = 0; j < n; j++)
      gemm('H', 'N', nao, nao, nao, cplx(-1.0), Acc.lesptr(j, n), X.data() + (size_t)j * es, ONE, r);
  }
  for (int n = j1; n <= j2; n++) {
    cplx *r = res + (size_t)(n - j1) * es;
    for (int j = n; j <= top; j++) gemm('N', 'N', nao, nao, nao, ONE, A.lesptr(n, j), X.data() + (size_t)j * es, ONE, r);
  }
}

// src/dyson_integrals.h:393-457
void Dyson::Cles3_tstp(int j1, int j2, const Green &A, const Green &Acc, const Green &B, const Green &Bcc, int m,
                       double beta, cplx *res) const {
  (void)Acc; (void)B;
  const int sig = B.sig;
  for (int j = 0; j <= ntau; j++) {
    const cplx *s = Bcc.tvptr(m, ntau - j);
    cplx *d = NTauTmp.data() + (size_t)j * es;
    for (int a = 0; a < nao; a++)
      for (int b = 0; b < nao; b++) d[a * nao + b] = (double)sig * I1 * std::conj(s[b * nao + a]);
  }
  for (int n = j1; n <= j2; n++) conv_lesser(res + (size_t)(n - j1) * es, A.tvptr(n, 0), NTauTmp.data(), beta);
}

// src/dyson_integrals.h:7-15 with the real kernel GMCT of src/greens.cpp:112-146
void Dyson::tv_it_conv(int m, int n, const Green &Sig, const Green &G, double beta, cplx *res) const {
  const cplx *K = mixing_kernel(G, beta);
  const int np = ntau + 1;
  std::vector<double> Kr(es);
  for (int i = 0; i < np; i++) {
    const cplx *Kmi = K + ((size_t)m * np + i) * es;
    for (int e = 0; e < es; e++) Kr[e] = Kmi[e].real();
    gemm('N', 'N', nao, nao, nao, ONE, Sig.tvptr(n, i), Kr.data(), ONE, res);
  }
}

// src/dyson_integrals.h:24-70
double Dyson::energy_conv(int tstp, const Green &Sig, const Green &G, double beta, double dt) const {
  int top = tstp >= k ? tstp : k, sig = G.sig;
  cplx res1 = 0, res2 = 0;
  auto trT = [&](const cplx *S, const cplx *Gm) {
    cplx s = 0;
    for (int a = 0; a < nao; a++)
      for (int b = 0; b < nao; b++) s += S[a * nao + b] * Gm[b * nao + a];
    return s;
  };
  for (int i = 0; i <= tstp; i++) res1 += I.gregory_weights(tstp, i) * trT(Sig.retptr(tstp, i), G.lesptr(i, tstp));
  for (int i = tstp + 1; i <= top; i++)
    res1 += std::conj(I.gregory_weights(tstp, i) * trT(Sig.retptr(i, tstp), G.lesptr(tstp, i)));
  for (int i = 0; i <= tstp; i++)
    res1 -= std::conj(I.gregory_weights(tstp, i) * trT(Sig.lesptr(i, tstp), G.retptr(tstp, i)));
  for (int i = tstp + 1; i <= top; i++) res1 -= I.gregory_weights(tstp, i) * trT(Sig.lesptr(tstp, i), G.retptr(i, tstp));
  res1 *= dt;
  res2 = conv_energy(Sig.tvptr(tstp, 0), G.tvptr(tstp, 0), beta, G.sig);
  res2 *= cplx(0., -1.);
  return (cplx(0., sig * 0.5) * (res1 + res2)).real();
}

// src/dyson_step.h:7-127
double Dyson::dyson_step_ret(int tstp, const Green &G, const Green &Sig, const cplx *hmf, double mu, double dt) const {
  double err = 0;
  const cplx ncplxi(0, -1);
  const int kn = k * nao;
  std::fill(M.begin(), M.end(), cplx(0));
  std::fill(Q.begin(), Q.begin() + (size_t)(tstp + 1) * es, cplx(0));
  cplx *Gtt = G.retptr(tstp, tstp);
  for (int e = 0; e < es; e++) Gtt[e] = 0;
  for (int a = 0; a < nao; a++) Gtt[a * nao + a] = ncplxi;
  auto Mblk = [&](int br, int bc, int a, int b) -> cplx & { return M[(size_t)(br * nao + a) * kn + bc * nao + b]; };
  for (int n = 1; n <= k; n++) {
    cplx *QB = Q.data() + (size_t)(n - 1) * es;
    for (int l = 0; l <= k; l++) {
      int bc = (l == 0) ? 0 : l - 1;
      if (l == 0) {
        cplx f = ncplxi / dt * I.poly_diff(n, l);
        for (int a = 0; a < nao; a++)
          for (int b = 0; b < nao; b++) QB[a * nao + b] += f * Gtt[b * nao + a];
      } else {
        cplx f = ncplxi / dt * I.poly_diff(n, l);
        for (int a = 0; a < nao; a++) Mblk(n - 1, bc, a, a) -= f;
      }
      if (l == n) {
        const cplx *h = hmf + (size_t)(tstp - n) * es;
        for (int a = 0; a < nao; a++)
          for (int b = 0; b < nao; b++) Mblk(n - 1, bc, a, b) += (a == b ? mu : 0.0) - h[b * nao + a];
      }
      if (l == 0) {
        gemm('N', 'N', nao, nao, nao, ONE, Gtt, Sig.retptr(tstp, tstp - n), ZERO, tmp.data());
        double w = dt * I.gregory_weights(n, l);
        for (int a = 0; a < nao; a++)
          for (int b = 0; b < nao; b++) QB[a * nao + b] += w * tmp[b * nao + a];
      } else {
        double w = dt * I.gregory_weights(n, l);
        if (tstp - l >= tstp - n) {
          const cplx *s = Sig.retptr(tstp - l, tstp - n);
          for (int a = 0; a < nao; a++)
            for (int b = 0; b < nao; b++) Mblk(n - 1, bc, a, b) -= w * s[b * nao + a];
        } else {
          const cplx *s = Sig.retptr(tstp - n, tstp - l);
          for (int a = 0; a < nao; a++)
            for (int b = 0; b < nao; b++) Mblk(n - 1, bc, a, b) += w * std::conj(s[a * nao + b]);
        }
      }
    }
  }
  std::vector<cplx> Xs(Q.begin(), Q.begin() + (size_t)k * es);
  fullpiv_solve(kn, M.data(), nao, Xs.data());
  for (int l = 0; l < k; l++) {
    cplx *g = G.retptr(tstp, tstp - l - 1);
    const cplx *x = Xs.data() + (size_t)l * es;
    double s = 0;
    for (int a = 0; a < nao; a++)
      for (int b = 0; b < nao; b++) s += std::norm(g[a * nao + b] - x[b * nao + a]);
    err += std::sqrt(s);
    for (int a = 0; a < nao; a++)
      for (int b = 0; b < nao; b++) g[a * nao + b] = x[b * nao + a];
  }
  for (int n = k + 1; n <= tstp; n++) {
    cplx *QB = Q.data() + (size_t)n * es;
    for (int l = 0; l <= k; l++)
      gemm('N', 'N', nao, nao, nao, I.gregory_weights(n, l), G.retptr(tstp, tstp - l), Sig.retptr(tstp - l, tstp - n), ONE, QB);
  }
  std::vector<cplx> MS(es), QT(es);
  for (int n = k + 1; n <= tstp; n++) {
    cplx *QB = Q.data() + (size_t)n * es;
    for (int e = 0; e < es; e++) QB[e] *= dt;
    for (int l = 1; l <= k + 1; l++) {
      cplx f = I.bd_weights(l) * ncplxi / dt;
      const cplx *g = G.retptr(tstp, tstp - n + l);
      for (int e = 0; e < es; e++) QB[e] += f * g[e];
    }
    const cplx *h = hmf + (size_t)(tstp - n) * es;
    const cplx *s = Sig.retptr(tstp - n, tstp - n);
    cplx d = mu - I.bd_weights(0) * ncplxi / dt;
    for (int a = 0; a < nao; a++)
      for (int b = 0; b < nao; b++)
        MS[a * nao + b] = -h[b * nao + a] - dt * I.omega(0) * s[b * nao + a] + (a == b ? d : 0.0);
    for (int a = 0; a < nao; a++)
      for (int b = 0; b < nao; b++) QT[a * nao + b] = QB[b * nao + a];
    fullpiv_solve(nao, MS.data(), nao, QT.data());
    cplx *g = G.retptr(tstp, tstp - n);
    double sum = 0;
    for (int a = 0; a < nao; a++)
      for (int b = 0; b < nao; b++) sum += std::norm(g[a * nao + b] - QT[b * nao + a]);
    err += std::sqrt(sum);
    for (int a = 0; a < nao; a++)
      for (int b = 0; b < nao; b++) g[a * nao + b] = QT[b * nao + a];
    for (int m = n + 1; m <= tstp; m++)
      gemm('N', 'N', nao, nao, nao, I.gregory_weights(m, n), g, Sig.retptr(tstp - n, tstp - m), ONE, Q.data() + (size_t)m * es);
  }
  return err;
}

// src/dyson_step.h:130-169
double Dyson::dyson_step_tv(int tstp, const Green &G, const Green &Sig, const cplx *hmf, double mu, double beta, double dt) const {
  double err = 0;
  const int np = ntau + 1;
  std::vector<cplx> old(G.tvptr(tstp, 0), G.tvptr(tstp, 0) + (size_t)np * es);
  std::memset((void *)G.tvptr(tstp, 0), 0, sizeof(cplx) * np * es);
  Ctv_tstp(tstp, G, Sig, Sig, G, G, beta, dt);
  cplx *gt = G.tvptr(tstp, 0);
  for (int l = 1; l <= k + 1; l++) {
    cplx f = -I1 / dt * I.bd_weights(l);
    const cplx *s = G.tvptr(tstp - l, 0);
    for (size_t e = 0; e < (size_t)np * es; e++) gt[e] += f * s[e];
  }
  std::vector<cplx> MM(es);
  const cplx *h = hmf + (size_t)tstp * es;
  const cplx *sr = Sig.retptr(tstp, tstp);
  cplx d = I1 / dt * I.bd_weights(0) + mu;
  for (int a = 0; a < nao; a++)
    for (int b = 0; b < nao; b++) MM[a * nao + b] = (a == b ? d : 0.0) - h[a * nao + b] - dt * I.omega(0) * sr[a * nao + b];
  // one factorisation, all right-hand sides (identical to per-column solves)
  std::vector<cplx> rhs((size_t)nao * nao * np);
  for (int m = 0; m < np; m++)
    for (int a = 0; a < nao; a++)
      for (int b = 0; b < nao; b++) rhs[(size_t)a * nao * np + m * nao + b] = gt[(size_t)m * es + a * nao + b];
  fullpiv_solve(nao, MM.data(), nao * np, rhs.data());
  for (int m = 0; m < np; m++) {
    double s = 0;
    for (int a = 0; a < nao; a++)
      for (int b = 0; b < nao; b++) {
        cplx x = rhs[(size_t)a * nao * np + m * nao + b];
        s += std::norm(old[(size_t)m * es + a * nao + b] - x);
        gt[(size_t)m * es + a * nao + b] = x;
      }
    err += std::sqrt(s);
  }
  return err;
}

// src/dyson_step.h:171-326
double Dyson::dyson_step_les(int n, const Green &G, const Green &Sig, const cplx *hmf, double mu, double beta, double dt) const {
  int num = n >= k ? n : k;
  double err = 0;
  const int kn = k * nao;
  {
    cplx *g = G.lesptr(0, n);
    const cplx *t = G.tvptr(n, 0);
    double s = 0;
    for (int a = 0; a < nao; a++)
      for (int b = 0; b < nao; b++) s += std::norm(g[a * nao + b] + std::conj(t[b * nao + a]));
    err += std::sqrt(s);
    for (int a = 0; a < nao; a++)
      for (int b = 0; b < nao; b++) g[a * nao + b] = -std::conj(t[b * nao + a]);
  }
  std::fill(Q.begin(), Q.begin() + (size_t)(num + 1) * es, cplx(0));
  std::fill(M.begin(), M.end(), cplx(0));
  Cles3_tstp(0, (n <= k ? k : n), Sig, Sig, G, G, n, beta, Q.data());
  Cles2_tstp(0, num, Sig, Sig, G, G, n, dt, Q.data());
  auto Mblk = [&](int br, int bc, int a, int b) -> cplx & { return M[(size_t)(br * nao + a) * kn + bc * nao + b]; };
  const cplx *G0n = G.lesptr(0, n);
  for (int m = 1; m <= k; m++) {
    cplx *QB = Q.data() + (size_t)m * es;
    for (int l = 0; l <= k; l++) {
      if (l == 0) {
        cplx f = I1 / dt * I.poly_diff(m, l);
        for (int e = 0; e < es; e++) QB[e] -= f * G0n[e];
      } else {
        cplx f = I1 / dt * I.poly_diff(m, l);
        for (int a = 0; a < nao; a++) Mblk(m - 1, l - 1, a, a) += f;
      }
      if (m == l) {
        const cplx *h = hmf + (size_t)l * es;
        for (int a = 0; a < nao; a++)
          for (int b = 0; b < nao; b++) Mblk(m - 1, l - 1, a, b) += (a == b ? mu : 0.0) - h[a * nao + b];
      }
      double w = dt * I.gregory_weights(m, l);
      if (l == 0) {
        gemm('N', 'N', nao, nao, nao, w, Sig.retptr(m, l), G0n, ONE, QB);
      } else if (m >= l) {
        const cplx *s = Sig.retptr(m, l);
        for (int a = 0; a < nao; a++)
          for (int b = 0; b < nao; b++) Mblk(m - 1, l - 1, a, b) -= w * s[a * nao + b];
      } else {
        const cplx *s = Sig.retptr(l, m);
        for (int a = 0; a < nao; a++)
          for (int b = 0; b < nao; b++) Mblk(m - 1, l - 1, a, b) += w * std::conj(s[b * nao + a]);
      }
    }
  }
  std::memcpy(X.data() + es, Q.data() + es, sizeof(cplx) * kn * nao);
  fullpiv_solve(kn, M.data(), nao, X.data() + es);
  {
    cplx *x0 = X.data();
    const cplx *t = G.tvptr(n, 0);
    for (int a = 0; a < nao; a++)
      for (int b = 0; b < nao; b++) x0[a * nao + b] = -std::conj(t[b * nao + a]);
  }
  std::vector<cplx> MS(es);
  for (int m = k + 1; m < n; m++) {
    cplx *QB = Q.data() + (size_t)m * es;
    const cplx *h = hmf + (size_t)m * es;
    const cplx *sr = Sig.retptr(m, m);
    cplx d = I1 / dt * I.bd_weights(0) + mu;
    for (int a = 0; a < nao; a++)
      for (int b = 0; b < nao; b++) MS[a * nao + b] = -h[a * nao + b] + (a == b ? d : 0.0) - dt * I.omega(0) * sr[a * nao + b];
    for (int l = 0; l < m; l++)
      gemm('N', 'N', nao, nao, nao, dt * I.gregory_weights(m, l), Sig.retptr(m, l), X.data() + (size_t)l * es, ONE, QB);
    for (int l = 1; l <= k + 1; l++) {
      cplx f = I1 / dt * I.bd_weights(l);
      const cplx *x = X.data() + (size_t)(m - l) * es;
      for (int e = 0; e < es; e++) QB[e] -= f * x[e];
    }
    std::memcpy(X.data() + (size_t)m * es, QB, sizeof(cplx) * es);
    fullpiv_solve(nao, MS.data(), nao, X.data() + (size_t)m * es);
  }
  cplx *Xn = X.data() + (size_t)n * es;
  std::fill(Xn, Xn + es, cplx(0));
  for (int jj = 0; jj <= k; jj++) {
    const cplx *g = G.lesptr(n - jj - 1, n - jj - 1);
    for (int e = 0; e < es; e++) Xn[e] += I.ex_weights(jj) * g[e];
  }
  cplx *QB = Q.data() + (size_t)n * es;
  for (int l = 0; l <= n; l++)
    gemm('N', 'N', nao, nao, nao, dt * I.gregory_weights(n, l), Sig.retptr(n, l), X.data() + (size_t)l * es, ONE, QB);
  {
    const cplx *h = hmf + (size_t)n * es;
    std::vector<cplx> hm(es);
    for (int a = 0; a < nao; a++)
      for (int b = 0; b < nao; b++) hm[a * nao + b] = h[a * nao + b] - (a == b ? mu : 0.0);
    gemm('N', 'N', nao, nao, nao, ONE, hm.data(), Xn, ONE, QB);
  }
  cplx *Q0 = Q.data();
  for (int a = 0; a < nao; a++)
    for (int b = 0; b < nao; b++) Q0[a * nao + b] = -I1 * (QB[a * nao + b] + std::conj(QB[b * nao + a]));
  for (int l = 1; l <= k + 1; l++) {
    double f = 1. / dt * I.bd_weights(l);
    const cplx *g = G.lesptr(n - l, n - l);
    for (int e = 0; e < es; e++) Q0[e] -= f * g[e];
  }
  MS.assign(es, cplx(0));
  for (int a = 0; a < nao; a++) MS[a * nao + a] = 1. / dt * I.bd_weights(0);
  std::memcpy(Xn, Q0, sizeof(cplx) * es);
  fullpiv_solve(nao, MS.data(), nao, Xn);
  for (int l = 0; l <= n; l++) {
    err += fro(G.lesptr(l, n), X.data() + (size_t)l * es, es);
    std::memcpy((void *)G.lesptr(l, n), X.data() + (size_t)l * es, sizeof(cplx) * es);
  }
  return err;
}

// src/dyson_step.h:329-382 (GF2 branch)
double Dyson::dyson_step(int n, const Green &G, const Green &Sig, const cplx *hmf, double mu, double beta, double dt) const {
  double err = 0;
  err += dyson_step_ret(n, G, Sig, hmf, mu, dt);
  err += dyson_step_tv(n, G, Sig, hmf, mu, beta, dt);
  err += dyson_step_les(n, G, Sig, hmf, mu, beta, dt);
  return err;
}

// src/dyson_boot.h:78-88 (the live TTI branch of dyson_start_ret)
double Dyson::dyson_start_ret(const Green &G, const Green &Sig, const cplx *hmf, double mu, double dt) const {
  (void)Sig; (void)hmf; (void)mu; (void)dt;
  double err = 0;
  std::vector<cplx> GR(es);
  for (int n = 0; n <= k; n++) {
    const cplx *a = G.tvptr(n, G.ntau), *b = G.tvptr(n, 0);
    for (int e = 0; e < es; e++) GR[e] = (double)G.sig * a[e] - b[e];
    if (n == 0) {
      std::fill(GR.begin(), GR.end(), cplx(0));
      for (int a2 = 0; a2 < nao; a2++) GR[a2 * nao + a2] = cplx(0., -1.);
    }
    for (int l = 0; l <= k - n; l++) {
      cplx *r = G.retptr(n + l, l);
      err += fro(GR.data(), r, es);
      std::memcpy(r, GR.data(), sizeof(cplx) * es);
    }
  }
  return err;
}

// src/dyson_boot.h:93-179
double Dyson::dyson_start_tv(const Green &G, const Green &Sig, const cplx *hmf, double mu, double beta, double dt) const {
  double err = 0;
  const int np = ntau + 1, kn = k * nao;
  for (int m = 0; m <= ntau; m++) {
    cplx *t = G.tvptr(0, m);
    const cplx *g = G.matptr(ntau - m);
    double s = 0;
    for (int e = 0; e < es; e++) {
      cplx v = (double)G.sig * I1 * g[e];
      s += std::norm(t[e] - v);
      t[e] = v;
    }
    err += std::sqrt(s);
  }
  // The k x k block matrix does not depend on m (it is rebuilt identically for
  // every m in the reference); factorise once and solve all m together.
  std::fill(M.begin(), M.end(), cplx(0));
  auto Mblk = [&](int br, int bc, int a, int b) -> cplx & { return M[(size_t)(br * nao + a) * kn + bc * nao + b]; };
  for (int n = 1; n <= k; n++)
    for (int l = 1; l <= k; l++) {
      cplx f = I1 * I.poly_diff(n, l) / dt;
      for (int a = 0; a < nao; a++) Mblk(n - 1, l - 1, a, a) += f;
      if (l == n) {
        const cplx *h = hmf + (size_t)l * es;
        for (int a = 0; a < nao; a++)
          for (int b = 0; b < nao; b++) Mblk(n - 1, l - 1, a, b) += (a == b ? mu : 0.0) - h[a * nao + b];
      }
      double w = dt * I.gregory_weights(n, l);
      if (n >= l) {
        const cplx *s = Sig.retptr(n, l);
        for (int a = 0; a < nao; a++)
          for (int b = 0; b < nao; b++) Mblk(n - 1, l - 1, a, b) -= w * s[a * nao + b];
      } else {
        const cplx *s = Sig.retptr(l, n);
        for (int a = 0; a < nao; a++)
          for (int b = 0; b < nao; b++) Mblk(n - 1, l - 1, a, b) += w * std::conj(s[b * nao + a]);
      }
    }
  // right-hand sides: rows (n-1, a), columns (m, b)
  std::vector<cplx> rhs((size_t)kn * nao * np, cplx(0));
  std::vector<cplx> q(es);
  for (int m = 0; m <= ntau; m++)
    for (int n = 1; n <= k; n++) {
      std::fill(q.begin(), q.end(), cplx(0));
      tv_it_conv(m, n, Sig, G, beta, q.data());
      cplx f = I1 * I.poly_diff(n, 0) / dt;
      const cplx *g0 = G.tvptr(0, m);
      for (int e = 0; e < es; e++) q[e] -= f * g0[e];
      gemm('N', 'N', nao, nao, nao, dt * I.gregory_weights(n, 0), Sig.retptr(n, 0), g0, ONE, q.data());
      for (int a = 0; a < nao; a++)
        for (int b = 0; b < nao; b++) rhs[(size_t)((n - 1) * nao + a) * nao * np + m * nao + b] = q[a * nao + b];
    }
  fullpiv_solve(kn, M.data(), nao * np, rhs.data());
  for (int m = 0; m <= ntau; m++)
    for (int l = 0; l < k; l++) {
      cplx *t = G.tvptr(l + 1, m);
      double s = 0;
      for (int a = 0; a < nao; a++)
        for (int b = 0; b < nao; b++) {
          cplx x = rhs[(size_t)(l * nao + a) * nao * np + m * nao + b];
          s += std::norm(t[a * nao + b] - x);
          t[a * nao + b] = x;
        }
      err += std::sqrt(s);
    }
  return err;
}

// src/dyson_boot.h:181-192
double Dyson::dyson_start_les(const Green &G, const Green &Sig, const cplx *hmf, double mu, double beta, double dt) const {
  (void)Sig; (void)hmf; (void)mu; (void)beta; (void)dt;
  double err = 0;
  for (int l = 0; l <= k; l++)
    for (int n = 0; n <= l; n++) {
      cplx *g = G.lesptr(n, l);
      const cplx *t = G.tvptr(l - n, 0);
      double s = 0;
      for (int a = 0; a < nao; a++)
        for (int b = 0; b < nao; b++) {
          cplx v = -std::conj(t[b * nao + a]);
          s += std::norm(g[a * nao + b] - v);
          g[a * nao + b] = v;
        }
      err += std::sqrt(s);
    }
  return err;
}

// src/dyson_boot.h:195-221
double Dyson::dyson_start(const Green &G, const Green &Sig, const cplx *hmf, double mu, double beta, double dt) const {
  double err = 0;
  err += dyson_start_tv(G, Sig, hmf, mu, beta, dt);
  err += dyson_start_ret(G, Sig, hmf, mu, dt);
  err += dyson_start_les(G, Sig, hmf, mu, beta, dt);
  return err;
}

// src/dyson_free.h:7-42
double fermi(double beta, double omega) {
  double arg = omega * beta;
  if (std::fabs(arg) > 100) return arg > 0.0 ? 0.0 : 1.0;
  return 1.0 / (1.0 + std::exp(arg));
}
double fermi_exp(double beta, double tau, double omega) {
  if (omega < 0) return std::exp(omega * tau) * fermi(beta, omega);
  return std::exp((tau - beta) * omega) * fermi(beta, -omega);
}

// src/dyson_free.h:226-297 (sig = -1)
void Dyson::G0_from_h0(const Green &G, double mu, const double *H0, double beta, double h) const {
  const int n = nao;
  std::vector<double> Hmu(es), ev(n), V(es);
  for (int a = 0; a < n; a++)
    for (int b = 0; b < n; b++) Hmu[a * n + b] = (a == b ? mu : 0.0) - H0[a * n + b];
  symeig(n, Hmu.data(), ev.data(), V.data());
  // Udt = exp(i h Hmu)
  std::vector<cplx> Udt(es, cplx(0)), U((size_t)(k + 1) * es, cplx(0));
  for (int a = 0; a < n; a++)
    for (int b = 0; b < n; b++) {
      cplx s = 0;
      for (int c = 0; c < n; c++) s += V[a * n + c] * std::exp(I1 * h * ev[c]) * V[b * n + c];
      Udt[a * n + b] = s;
    }
  for (int a = 0; a < n; a++) U[a * n + a] = 1;
  for (int t = 1; t <= k; t++)
    gemm('N', 'N', n, n, n, ONE, U.data() + (size_t)(t - 1) * es, Udt.data(), ZERO, U.data() + (size_t)t * es);
  std::vector<cplx> F(es), tmpm(es);
  for (int m = 0; m <= ntau; m++) {
    double tau = L.tau(m, beta);
    for (int a = 0; a < n; a++)
      for (int b = 0; b < n; b++) {
        cplx s = 0;
        for (int c = 0; c < n; c++) s += V[a * n + c] * fermi_exp(beta, tau, -ev[c]) * V[b * n + c];
        F[a * n + b] = s;
      }
    for (int t = 0; t <= k; t++) gemm('N', 'N', n, n, n, I1, U.data() + (size_t)t * es, F.data(), ZERO, G.tvptr(t, m));
  }
  std::vector<cplx> val(es);
  for (int a = 0; a < n; a++)
    for (int b = 0; b < n; b++) {
      cplx s = 0;
      for (int c = 0; c < n; c++) s += V[a * n + c] * fermi(beta, -ev[c]) * V[b * n + c];
      val[a * n + b] = s;
    }
  for (int m = 0; m <= k; m++)
    for (int t = 0; t <= m; t++) {
      gemm('N', 'H', n, n, n, cplx(0, -1), U.data() + (size_t)m * es, U.data() + (size_t)t * es, ZERO, G.retptr(m, t));
      gemm('N', 'N', n, n, n, I1, U.data() + (size_t)t * es, val.data(), ZERO, tmpm.data());
      gemm('N', 'H', n, n, n, ONE, tmpm.data(), U.data() + (size_t)m * es, ZERO, G.lesptr(t, m));
    }
}

// src/dyson_free.h:161-176 (Matsubara branch of the non-constant version)
void Dyson::G0_matsubara(const Green &G, double mu, const double *H0, double beta) const {
  const int n = nao;
  std::vector<double> Hmu(es), ev(n), V(es);
  for (int a = 0; a < n; a++)
    for (int b = 0; b < n; b++) Hmu[a * n + b] = (a == b ? mu : 0.0) - H0[a * n + b];
  symeig(n, Hmu.data(), ev.data(), V.data());
  for (int m = 0; m <= ntau; m++) {
    double tau = L.tau(m, beta);
    cplx *g = G.matptr(m);
    for (int a = 0; a < n; a++)
      for (int b = 0; b < n; b++) {
        double s = 0;
        for (int c = 0; c < n; c++) s += V[a * n + c] * fermi_exp(beta, tau, ev[c]) * V[b * n + c];
        g[a * n + b] = -s;
      }
  }
}

}  // namespace orc
