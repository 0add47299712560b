// Initial-state helpers on the GPU: the non-interacting Green's function of a
// constant mean field (dyson::G0_from_h0, src/dyson_free.h:226-297, for any
// number of time steps) and the time-translation-invariant continuation of one
// slice (the relation the bootstrap imposes on G^R, src/dyson_boot.h:78-88).
#include "dyson.h"
#include "quadrature.h"
#include <cmath>
#include <vector>

namespace ned {

namespace {

// out = pref * V diag(c) V^T for every block; c_k given per block
__global__ void kern_vdv(const double *V, const z2 *coef, int n, int nblk, z2 pref, z2 *out) {
  const size_t es = (size_t)n * n;
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)nblk * es) return;
  int b = (int)(idx / es), e = (int)(idx % es), i = e / n, j = e % n;
  const z2 *c = coef + (size_t)b * n;
  z2 s = {0.0, 0.0};
  for (int k = 0; k < n; k++) {
    double v = V[i * n + k] * V[j * n + k];
    s.x += v * c[k].x;
    s.y += v * c[k].y;
  }
  out[idx] = zmul(pref, s);
}

// ret(t, t') = R[t - t'], les(t', t) = L[t - t'] for 0 <= t' <= t <= tmax
__global__ void kern_fill_tti(const z2 *R, const z2 *L, int n, int tmax, z2 *ret, z2 *les) {
  const size_t es = (size_t)n * n;
  const long long nb = (long long)(tmax + 1) * (tmax + 2) / 2;
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nb * (long long)es) return;
  long long b = idx / es;
  int e = (int)(idx % es);
  // packed block b = t(t+1)/2 + s
  int t = (int)((sqrt(8.0 * (double)b + 1.0) - 1.0) / 2.0);
  while ((long long)(t + 1) * (t + 2) / 2 <= b) t++;
  while ((long long)t * (t + 1) / 2 > b) t--;
  int s = (int)(b - (long long)t * (t + 1) / 2);
  // ret block (t, s): difference t - s;  les block (s, t): difference t - s
  ret[idx] = R[(size_t)(t - s) * es + e];
  les[idx] = L[(size_t)(t - s) * es + e];
}

// the same for rows t0..tmax of a ring of G rows (row t in slot t mod gring,
// rowlen blocks a slot): one thread per element of (row, column s <= tmax)
__global__ void kern_fill_tti_ring(const z2 *R, const z2 *L, int n, int t0, int tmax, int gring, long long rowlen,
                                   z2 *ret, z2 *les) {
  const size_t es = (size_t)n * n;
  const long long per = (long long)(tmax + 1) * (long long)es;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)(tmax - t0 + 1) * per) return;
  const int t = t0 + (int)(idx / per);
  const long long r = idx % per;
  const int s = (int)(r / es), e = (int)(r % es);
  if (s > t) return;
  const size_t o = ((size_t)(t % gring) * (size_t)rowlen + s) * es + e;
  ret[o] = R[(size_t)(t - s) * es + e];
  les[o] = L[(size_t)(t - s) * es + e];
}

// copy slice tsrc to every t <= tmax (tv), and the triangles by difference
// (SE: the triangles' element type, complex FP64 or FP32)
template <typename SE>
__global__ void kern_tti_extend(SE *ret, SE *les, z2 *tv, int n, int np, int tsrc, int tmax) {
  const size_t es = (size_t)n * n;
  const long long nb = (long long)(tmax + 1) * (tmax + 2) / 2;
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long ntri = nb * (long long)es;
  if (idx < ntri) {
    long long b = idx / es;
    int e = (int)(idx % es);
    int t = (int)((sqrt(8.0 * (double)b + 1.0) - 1.0) / 2.0);
    while ((long long)(t + 1) * (t + 2) / 2 <= b) t++;
    while ((long long)t * (t + 1) / 2 > b) t--;
    int s = (int)(b - (long long)t * (t + 1) / 2);
    if (t == tsrc) return;
    const size_t rowS = (size_t)tsrc * (tsrc + 1) / 2;
    ret[idx] = ret[(rowS + tsrc - (t - s)) * es + e];
    les[idx] = les[(rowS + tsrc - (t - s)) * es + e];
    return;
  }
  idx -= ntri;
  if (idx >= (long long)(tmax + 1) * np * (long long)es) return;
  int t = (int)(idx / ((long long)np * es));
  if (t == tsrc) return;
  long long r = idx % ((long long)np * es);
  tv[idx] = tv[(size_t)tsrc * np * es + r];
}

}  // namespace

void g0_from_h0(Ctx &c, int tmax, const double *fock, double mu, double beta, double dt) {
  require(tmax >= 0 && tmax <= c.nt, "g0_from_h0: tmax out of range");
  const int n = c.nao, np = c.np;
  const size_t es = c.es;
  std::vector<double> Hmu(es), lam(n), V(es);
  for (int a = 0; a < n; a++)
    for (int b = 0; b < n; b++) Hmu[a * n + b] = (a == b ? mu : 0.0) - fock[a * n + b];
  sym_eig(n, Hmu.data(), lam.data(), V.data());
  auto fermi = [&](double w) {
    double a = w * beta;
    if (std::fabs(a) > 100) return a > 0 ? 0.0 : 1.0;
    return 1.0 / (1.0 + std::exp(a));
  };
  auto fexp = [&](double tau, double w) {
    return w < 0 ? std::exp(w * tau) * fermi(w) : std::exp((tau - beta) * w) * fermi(-w);
  };
  GLL g(c.ntau);
  // coefficient tables: R[d]_k = e^{i lam_k d dt}; L[d]_k = e^{-i lam_k d dt} f(-lam_k);
  // tv[t][m]_k = e^{i lam_k t dt} fexp(tau_m, -lam_k)
  std::vector<z2> cR((size_t)(tmax + 1) * n), cL((size_t)(tmax + 1) * n), cT((size_t)(tmax + 1) * np * n);
  for (int d = 0; d <= tmax; d++)
    for (int k = 0; k < n; k++) {
      double ph = lam[k] * d * dt, oc = fermi(-lam[k]);
      cR[(size_t)d * n + k] = {std::cos(ph), std::sin(ph)};
      cL[(size_t)d * n + k] = {std::cos(ph) * oc, -std::sin(ph) * oc};
      for (int m = 0; m < np; m++) {
        double f = fexp(beta * (g.x[m] + 1) / 2, -lam[k]);
        cT[((size_t)d * np + m) * n + k] = {std::cos(ph) * f, std::sin(ph) * f};
      }
    }
  DevBuf dV, dc, dR, dL;
  dV.ensure(es * sizeof(double));
  dc.ensure(std::max(cT.size(), cR.size()) * sizeof(z2));
  dR.ensure((size_t)(tmax + 1) * es * sizeof(z2));
  dL.ensure((size_t)(tmax + 1) * es * sizeof(z2));
  hipStream_t s = c.stream;
  NED_HIP(hipMemcpyAsync(dV.p, V.data(), es * sizeof(double), hipMemcpyHostToDevice, s));
  auto nb = [](long long x) { return (unsigned)((x + 255) / 256); };
  NED_HIP(hipMemcpyAsync(dc.p, cR.data(), cR.size() * sizeof(z2), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(kern_vdv, dim3(nb((long long)(tmax + 1) * es)), dim3(256), 0, s, dV.as<double>(), dc.as<z2>(), n,
                     tmax + 1, z2{0.0, -1.0}, dR.as<z2>());
  NED_HIP(hipStreamSynchronize(s));
  NED_HIP(hipMemcpyAsync(dc.p, cL.data(), cL.size() * sizeof(z2), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(kern_vdv, dim3(nb((long long)(tmax + 1) * es)), dim3(256), 0, s, dV.as<double>(), dc.as<z2>(), n,
                     tmax + 1, z2{0.0, 1.0}, dL.as<z2>());
  NED_HIP(hipStreamSynchronize(s));
  if (c.gring) {  // only the rows the ring holds
    const int t0 = std::max(0, tmax - c.gring + 1);
    const long long tot = (long long)(tmax - t0 + 1) * (tmax + 1) * (long long)es;
    hipLaunchKernelGGL(kern_fill_tti_ring, dim3(nb(tot)), dim3(256), 0, s, dR.as<z2>(), dL.as<z2>(), n, t0, tmax,
                       c.gring, (long long)c.nt + 1, c.G.ret.as<z2>(), c.G.les.as<z2>());
  } else {
    long long ntri = (long long)(tmax + 1) * (tmax + 2) / 2;
    hipLaunchKernelGGL(kern_fill_tti, dim3(nb(ntri * (long long)es)), dim3(256), 0, s, dR.as<z2>(), dL.as<z2>(), n,
                       tmax, c.G.ret.as<z2>(), c.G.les.as<z2>());
  }
  NED_HIP(hipMemcpyAsync(dc.p, cT.data(), cT.size() * sizeof(z2), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(kern_vdv, dim3(nb((long long)(tmax + 1) * np * es)), dim3(256), 0, s, dV.as<double>(), dc.as<z2>(),
                     n, (tmax + 1) * np, z2{0.0, 1.0}, c.G.tv.as<z2>());
  check_launch("g0_from_h0");
  NED_HIP(hipStreamSynchronize(s));
  c.g_top = tmax;
  invalidate_minv(c);
}

void tti_extend(Ctx &c, int which, int tsrc) {
  require(tsrc >= 0 && tsrc <= c.nt, "tti_extend: tsrc out of range");
  require(which == 1 || !c.gring, "tti_extend: G keeps only its last rows (ring storage)");
  Green &g = which == 1 ? c.S : c.G;
  const size_t es = c.es;
  long long ntri = (long long)(tsrc + 1) * (tsrc + 2) / 2;
  long long tot = ntri * (long long)es + (long long)(tsrc + 1) * c.np * (long long)es;
  const dim3 grid((unsigned)((tot + 255) / 256));
  if (which == 1 && c.s32)
    hipLaunchKernelGGL(kern_tti_extend<float2>, grid, dim3(256), 0, c.stream, g.ret.as<float2>(), g.les.as<float2>(),
                       g.tv.as<z2>(), c.nao, c.np, tsrc, tsrc);
  else
    hipLaunchKernelGGL(kern_tti_extend<z2>, grid, dim3(256), 0, c.stream, g.ret.as<z2>(), g.les.as<z2>(),
                       g.tv.as<z2>(), c.nao, c.np, tsrc, tsrc);
  check_launch("tti_extend");
  NED_HIP(hipStreamSynchronize(c.stream));
  invalidate_minv(c);
}

}  // namespace ned
