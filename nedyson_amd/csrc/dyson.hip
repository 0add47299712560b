// Dyson timestep on the GPU in FP64: dyson::dyson_step (src/dyson_step.h:7-382)
// with its history integrals (src/dyson_integrals.h), the predictor
// dyson::Extrapolate (src/dyson.cpp:47-117), the HF mean field
// (src/molHFSolver.cpp:81-110), the bootstrap dyson_start
// (src/dyson_boot.h:6-221) and the energy (src/dyson_integrals.h:24-70).
//
// Volterra rows (ret) and columns (les) are solved right-looking: once the
// block X_n is known, one launch adds its contribution to the history sum of
// every later n' in parallel (one workgroup per n'), and the workgroup that
// owns n+1 completes that block in the same launch.  The per-time matrices
//   M_j = (mu + i bd_0/dt) 1 - hmf(j) - dt omega_0 Sigma^R(j,j)
// that the reference LU-factorises inside both recurrences
// (src/dyson_step.h:95-103, 259-278) depend only on the past time j, so their
// inverses are computed once, in one batched launch, and cached.
#include "dyson.h"
#include "blockops.h"
#include "quadrature.h"
#include <cmath>
#include <vector>
#include <algorithm>

namespace ned {

WTab make_wtab(int k) {
  require(k >= 1 && k <= KMAX, "integrator order k out of range (1..8)");
  Weights W(k);
  WTab t{};
  t.k = k;
  const int k1 = k + 1;
  for (int i = 0; i < k1 * k1; i++) { t.start[i] = W.start[i]; t.Omega[i] = W.Omega[i]; t.pdiff[i] = W.diff[i]; }
  for (int i = 0; i < k + 2; i++) t.bd[i] = W.bd[i];
  for (int i = 0; i < k1; i++) t.ex[i] = W.ex[i];
  return t;
}

// the integrator tables in device memory (kernels index them at run time, so
// they are read through a pointer rather than copied per thread from the
// kernel arguments)
const WTab *dev_wtab(Ctx &c) {
  if (c.wtab_k != c.k) {
    const WTab w = make_wtab(c.k);
    c.wtab_d.ensure(sizeof(WTab));
    NED_HIP(hipMemcpyAsync(c.wtab_d.p, &w, sizeof(WTab), hipMemcpyHostToDevice, c.stream));
    NED_HIP(hipStreamSynchronize(c.stream));
    c.wtab_k = c.k;
  }
  return (const WTab *)c.wtab_d.p;
}

namespace {

__device__ __forceinline__ size_t dtri(int t) { return (size_t)t * (t + 1) / 2; }

// gregory_weights(n, j), src/integration.h:78-91
__device__ __forceinline__ double greg(const WTab &w, int n, int j) {
  const int k = w.k, k1 = k + 1;
  const double *om = w.Omega + k * k1;
  if (n <= k && j <= k) return w.start[n * k1 + j];
  if (n <= 2 * k + 1) return j <= k ? w.Omega[(n - k - 1) * k1 + j] : om[n - j];
  if (j <= k) return om[j];
  if (j < n - k) return 1.0;
  return om[n - j];
}

// One system's operands of a Dyson step.  SE is the element type of the
// stored Sigma^R / Sigma^< triangles: complex FP64 (z2), or complex FP32
// (float2) for a context that keeps its self-energy history in FP32; every
// read of them goes through ldz(), which widens to FP64.  The G^R / G^<
// triangles are either whole (gring = 0, row i at block i(i+1)/2) or a ring
// of the last gring rows (row i in slot i mod gring, growlen blocks a slot).
template <typename SE>
struct PtT {
  z2 *Gret, *Gles, *Gtv;
  const z2 *Gmat;
  const SE *Sret, *Sles;
  const z2 *Stv;
  const z2 *hmf;
  int n, np, T, sig;
  int gring;
  long long growlen;
  double mu, beta, dt;
  __device__ size_t es() const { return (size_t)n * n; }
  __device__ size_t grow(int i) const { return gring ? (size_t)(i % gring) * (size_t)growlen : dtri(i); }
  __device__ z2 *gr(int i, int j) const { return Gret + (grow(i) + j) * es(); }
  __device__ z2 *gl(int i, int j) const { return Gles + (grow(j) + i) * es(); }
  __device__ z2 *gtv(int i, int m) const { return Gtv + ((size_t)i * np + m) * es(); }
  __device__ const SE *sr(int i, int j) const { return Sret + (dtri(i) + j) * es(); }
  __device__ const SE *sl(int i, int j) const { return Sles + (dtri(j) + i) * es(); }
  __device__ const z2 *stv(int i, int m) const { return Stv + ((size_t)i * np + m) * es(); }
  __device__ const z2 *h(int t) const { return hmf + (size_t)t * es(); }
};
using Pt = PtT<z2>;
using Pt32 = PtT<float2>;

// the same operands with the Sigma triangles read as FP32 (the layouts of the
// two instances are identical: only the pointee type of Sret / Sles differs)
inline Pt32 as_s32(const Pt &p) {
  Pt32 q;
  q.Gret = p.Gret; q.Gles = p.Gles; q.Gtv = p.Gtv; q.Gmat = p.Gmat;
  q.Sret = reinterpret_cast<const float2 *>(p.Sret);
  q.Sles = reinterpret_cast<const float2 *>(p.Sles);
  q.Stv = p.Stv; q.hmf = p.hmf;
  q.n = p.n; q.np = p.np; q.T = p.T; q.sig = p.sig;
  q.gring = p.gring; q.growlen = p.growlen;
  q.mu = p.mu; q.beta = p.beta; q.dt = p.dt;
  return q;
}

// One system's operands of the dependency chains and bulk updates of the
// Dyson recurrences, for the grouped launches that step many systems at once:
// the two accumulators (QA the chain's, Q the bulk updates'), the lesser
// column X, the cached inverses and the retarded chain's band norms.
template <typename SE>
struct ChainArgT {
  PtT<SE> p;
  z2 *QA, *Q, *X;
  const z2 *Minv;
  double *d2;
  int nb;  // 16-row / 16-column bands of a block
};
using ChainArg = ChainArgT<z2>;


// ------------------------------------------------------------ batched solve
// A_s X = B_s for every system s (N x N, R right-hand sides), Gaussian
// elimination with partial pivoting, one workgroup per system, in place in
// global memory (A destroyed, B <- X).  The reference uses Eigen::FullPivLU;
// on these diagonally dominant systems both agree to rounding.
__global__ void __launch_bounds__(256) kern_gesv(z2 *A, long long sA, z2 *B, long long sB, int N, int R, int *info) {
  z2 *a = A + (size_t)blockIdx.x * sA, *b = B + (size_t)blockIdx.x * sB;
  __shared__ double rv[256];
  __shared__ int ri[256];
  __shared__ z2 f[320];
  const int tid = threadIdx.x;
  for (int c = 0; c < N; c++) {
    double best = -1; int bi = c;
    for (int r = c + tid; r < N; r += 256) {
      z2 v = a[(size_t)r * N + c];
      double m = fabs(v.x) + fabs(v.y);
      if (m > best) { best = m; bi = r; }
    }
    rv[tid] = best; ri[tid] = bi;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (tid < w && rv[tid + w] > rv[tid]) { rv[tid] = rv[tid + w]; ri[tid] = ri[tid + w]; }
      __syncthreads();
    }
    const int p = ri[0];
    if (rv[0] <= 0.0) { if (tid == 0) atomicAdd(info, 1); }
    if (p != c) {
      for (int j = tid; j < N; j += 256) { z2 t = a[(size_t)c * N + j]; a[(size_t)c * N + j] = a[(size_t)p * N + j]; a[(size_t)p * N + j] = t; }
      for (int j = tid; j < R; j += 256) { z2 t = b[(size_t)c * R + j]; b[(size_t)c * R + j] = b[(size_t)p * R + j]; b[(size_t)p * R + j] = t; }
    }
    __syncthreads();
    const z2 piv = a[(size_t)c * N + c];
    const double den = piv.x * piv.x + piv.y * piv.y;
    const z2 inv = den > 0 ? z2{piv.x / den, -piv.y / den} : z2{0.0, 0.0};
    for (int r = c + 1 + tid; r < N; r += 256) f[r] = zmul(a[(size_t)r * N + c], inv);
    __syncthreads();
    const int rows = N - c - 1, colsA = N - c - 1;
    for (long long e = tid; e < (long long)rows * (colsA + R); e += 256) {
      int r = c + 1 + (int)(e / (colsA + R));
      int j = (int)(e % (colsA + R));
      z2 fr = f[r];
      if (j < colsA) {
        int jj = c + 1 + j;
        a[(size_t)r * N + jj] = zsub(a[(size_t)r * N + jj], zmul(fr, a[(size_t)c * N + jj]));
      } else {
        int jj = j - colsA;
        b[(size_t)r * R + jj] = zsub(b[(size_t)r * R + jj], zmul(fr, b[(size_t)c * R + jj]));
      }
    }
    __syncthreads();
  }
  for (int c = N - 1; c >= 0; c--) {
    const z2 piv = a[(size_t)c * N + c];
    const double den = piv.x * piv.x + piv.y * piv.y;
    const z2 inv = den > 0 ? z2{piv.x / den, -piv.y / den} : z2{0.0, 0.0};
    for (int j = tid; j < R; j += 256) b[(size_t)c * R + j] = zmul(b[(size_t)c * R + j], inv);
    __syncthreads();
    for (long long e = tid; e < (long long)c * R; e += 256) {
      int r = (int)(e / R), j = (int)(e % R);
      b[(size_t)r * R + j] = zsub(b[(size_t)r * R + j], zmul(a[(size_t)r * N + c], b[(size_t)c * R + j]));
    }
    __syncthreads();
  }
}

// The same elimination for N, R <= 64 with both matrices resident in LDS
// (one workgroup per system).
constexpr int LDS_N = 64;
__global__ void __launch_bounds__(256) kern_gesv_lds(z2 *A, long long sA, z2 *B, long long sB, int N, int R, int *info) {
  __shared__ z2 a[LDS_N][LDS_N + 1];
  __shared__ z2 b[LDS_N][LDS_N + 1];
  __shared__ double rv[64];
  __shared__ int ri[64];
  __shared__ z2 f[LDS_N];
  z2 *ga = A + (size_t)blockIdx.x * sA, *gb = B + (size_t)blockIdx.x * sB;
  const int tid = threadIdx.x;
  for (int e = tid; e < N * N; e += 256) a[e / N][e % N] = ga[e];
  for (int e = tid; e < N * R; e += 256) b[e / R][e % R] = gb[e];
  __syncthreads();
  for (int c = 0; c < N; c++) {
    if (tid < 64) {
      double best = -1; int bi = c;
      const int r = c + tid;
      if (r < N) { best = fabs(a[r][c].x) + fabs(a[r][c].y); bi = r; }
      rv[tid] = best; ri[tid] = bi;
    }
    __syncthreads();
    for (int w = 32; w > 0; w >>= 1) {
      if (tid < w && rv[tid + w] > rv[tid]) { rv[tid] = rv[tid + w]; ri[tid] = ri[tid + w]; }
      __syncthreads();
    }
    const int p = ri[0];
    if (rv[0] <= 0.0 && tid == 0) atomicAdd(info, 1);
    if (p != c) {
      for (int j = tid; j < N; j += 256) { z2 t = a[c][j]; a[c][j] = a[p][j]; a[p][j] = t; }
      for (int j = tid; j < R; j += 256) { z2 t = b[c][j]; b[c][j] = b[p][j]; b[p][j] = t; }
    }
    __syncthreads();
    const z2 piv = a[c][c];
    const double den = piv.x * piv.x + piv.y * piv.y;
    const z2 inv = den > 0 ? z2{piv.x / den, -piv.y / den} : z2{0.0, 0.0};
    for (int r = c + 1 + tid; r < N; r += 256) f[r] = zmul(a[r][c], inv);
    __syncthreads();
    const int rows = N - c - 1, colsA = N - c - 1, w = colsA + R;
    for (int e = tid; e < rows * w; e += 256) {
      const int r = c + 1 + e / w, j = e % w;
      const z2 fr = f[r];
      if (j < colsA) a[r][c + 1 + j] = zsub(a[r][c + 1 + j], zmul(fr, a[c][c + 1 + j]));
      else b[r][j - colsA] = zsub(b[r][j - colsA], zmul(fr, b[c][j - colsA]));
    }
    __syncthreads();
  }
  for (int c = N - 1; c >= 0; c--) {
    const z2 piv = a[c][c];
    const double den = piv.x * piv.x + piv.y * piv.y;
    const z2 inv = den > 0 ? z2{piv.x / den, -piv.y / den} : z2{0.0, 0.0};
    for (int j = tid; j < R; j += 256) b[c][j] = zmul(b[c][j], inv);
    __syncthreads();
    for (int e = tid; e < c * R; e += 256) {
      const int r = e / R, j = e % R;
      b[r][j] = zsub(b[r][j], zmul(a[r][c], b[c][j]));
    }
    __syncthreads();
  }
  for (int e = tid; e < N * R; e += 256) gb[e] = b[e / R][e % R];
}

// Gauss-Jordan with partial pivoting for one large system, one column per
// pair of launches: the pivot search and row swap (one workgroup), then the
// elimination of that column from every other row across the whole chip.
// Columns left of the pivot are never read again and are not updated.
__global__ void __launch_bounds__(256) kern_gj_pivot(z2 *A, z2 *B, int N, int R, int c, z2 *pinv, int *info) {
  __shared__ double rv[256];
  __shared__ int ri[256];
  const int tid = threadIdx.x;
  double best = -1; int bi = c;
  for (int r = c + tid; r < N; r += 256) {
    const z2 v = A[(size_t)r * N + c];
    const double m = fabs(v.x) + fabs(v.y);
    if (m > best) { best = m; bi = r; }
  }
  rv[tid] = best; ri[tid] = bi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w && rv[tid + w] > rv[tid]) { rv[tid] = rv[tid + w]; ri[tid] = ri[tid + w]; }
    __syncthreads();
  }
  const int p = ri[0];
  if (rv[0] <= 0.0 && tid == 0) atomicAdd(info, 1);
  if (p != c) {
    for (int j = c + tid; j < N; j += 256) {
      z2 t = A[(size_t)c * N + j]; A[(size_t)c * N + j] = A[(size_t)p * N + j]; A[(size_t)p * N + j] = t;
    }
    for (int j = tid; j < R; j += 256) {
      z2 t = B[(size_t)c * R + j]; B[(size_t)c * R + j] = B[(size_t)p * R + j]; B[(size_t)p * R + j] = t;
    }
  }
  __syncthreads();
  if (tid == 0) {
    const z2 piv = A[(size_t)c * N + c];
    const double den = piv.x * piv.x + piv.y * piv.y;
    pinv[c] = den > 0 ? z2{piv.x / den, -piv.y / den} : z2{0.0, 0.0};
  }
}

__global__ void kern_gj_elim(z2 *A, z2 *B, int N, int R, int c, const z2 *pinv) {
  const int wA = N - c - 1, w = wA + R;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)(N - 1) * w) return;
  int r = (int)(e / w);
  if (r >= c) r++;  // every row but the pivot row
  const int j = (int)(e % w);
  const z2 f = zmul(A[(size_t)r * N + c], pinv[c]);
  if (j < wA) {
    const int jj = c + 1 + j;
    A[(size_t)r * N + jj] = zsub(A[(size_t)r * N + jj], zmul(f, A[(size_t)c * N + jj]));
  } else {
    const int jj = j - wA;
    B[(size_t)r * R + jj] = zsub(B[(size_t)r * R + jj], zmul(f, B[(size_t)c * R + jj]));
  }
}

__global__ void kern_gj_scale(z2 *B, int N, int R, const z2 *pinv) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)N * R) return;
  B[e] = zmul(B[e], pinv[e / R]);
}

// M_j for the listed j into A (and identity into B), for inversion
template <typename SE>
__global__ void kern_build_mj(PtT<SE> p, const int *js, int cnt, double bd0, double om0, z2 *A, z2 *B) {
  const size_t es = p.es();
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)cnt * es) return;
  int s = (int)(idx / es), e = (int)(idx % es), a = e / p.n, b = e % p.n;
  int j = js[s];
  z2 hv = p.h(j)[e], sv = ldz(p.sr(j, j) + (e));
  z2 v = {-hv.x - p.dt * om0 * sv.x, -hv.y - p.dt * om0 * sv.y};
  if (a == b) { v.x += p.mu; v.y += bd0 / p.dt; }
  A[(size_t)s * es + e] = v;
  B[(size_t)s * es + e] = (a == b) ? z2{1.0, 0.0} : z2{0.0, 0.0};
}

__global__ void kern_scatter_blocks(const z2 *src, const int *js, int cnt, size_t es, z2 *dst) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)cnt * es) return;
  int s = (int)(idx / es);
  size_t e = idx % es;
  dst[(size_t)js[s] * es + e] = src[idx];
}

// ------------------------------------------------------------ retarded
// k x k block start-up system of src/dyson_step.h:28-65: build_M writes M
// (kn x kn, depends on times T-k..T-1 only) and sets B to the identity;
// otherwise the kernel writes Q (kn x n) and sets G^R(T,T) = -i.
template <typename SE>
__global__ void __launch_bounds__(256) kern_ret_start(PtT<SE> p, const WTab *__restrict__ wtab, int build_M, z2 *M, z2 *Q) {
  const WTab &w = *wtab;
  const int n = p.n, k = w.k, kn = k * n, T = p.T, k1 = k + 1;
  const long long gid = (long long)blockIdx.x * 256 + threadIdx.x, gsz = (long long)gridDim.x * 256;
  if (build_M) {
    for (long long e = gid; e < (long long)kn * kn; e += gsz) {
      int row = (int)(e / kn), col = (int)(e % kn);
      int nn = row / n + 1, a = row % n, l = col / n + 1, b = col % n;
      z2 v = {0.0, 0.0};
      // -(-i/dt) pd(n,l) I = (i/dt) pd I
      if (a == b) v.y += w.pdiff[nn * k1 + l] / p.dt;
      if (l == nn) {
        z2 hv = p.h(T - nn)[b * n + a];
        v.x -= hv.x; v.y -= hv.y;
        if (a == b) v.x += p.mu;
      }
      double wt = p.dt * greg(w, nn, l);
      if (l <= nn) {
        z2 s = ldz(p.sr(T - l, T - nn) + (b * n + a));
        v.x -= wt * s.x; v.y -= wt * s.y;
      } else {
        z2 s = ldz(p.sr(T - nn, T - l) + (a * n + b));
        v.x += wt * s.x; v.y -= wt * s.y;
      }
      M[e] = v;
      Q[e] = (row == col) ? z2{1.0, 0.0} : z2{0.0, 0.0};
    }
    return;
  }
  // G^R(T,T) = -i 1 (src/dyson_step.h:29)
  z2 *Gtt = p.gr(T, T);
  for (long long e = gid; e < (long long)n * n; e += gsz) Gtt[e] = (e / n == e % n) ? z2{0.0, -1.0} : z2{0.0, 0.0};
  // Q: (-i/dt) pd(n,0) G(T,T)^T + dt w(n,0) (G(T,T) Sigma^R(T,T-n))^T, with
  // G(T,T) = -i 1: the product is -i Sigma^R(T,T-n)^T
  for (long long e = gid; e < (long long)kn * n; e += gsz) {
    int row = (int)(e / n), b = (int)(e % n);
    int nn = row / n + 1, a = row % n;
    const z2 g = (a == b) ? z2{0.0, -1.0} : z2{0.0, 0.0};
    double pd = w.pdiff[nn * k1 + 0] / p.dt;
    z2 v = {pd * g.y, -pd * g.x};  // -i * pd * g
    const z2 s = zmul(z2{0.0, -1.0}, ldz(p.sr(T, T - nn) + (b * n + a)));
    double wt = p.dt * greg(w, nn, 0);
    v.x += wt * s.x; v.y += wt * s.y;
    Q[e] = v;
  }
}

// C (R x Cc) = A (R x R) B (R x Cc), z2 row-major; one thread per output
__global__ void kern_zmm(const z2 *A, const z2 *B, int R, int Cc, z2 *C) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)R * Cc) return;
  int r = (int)(idx / Cc), c = (int)(idx % Cc);
  z2 s = {0.0, 0.0};
  for (int q = 0; q < R; q++) s = zadd(s, zmul(A[(size_t)r * R + q], B[(size_t)q * Cc + c]));
  C[idx] = s;
}

// G^R(T,T-l-1) = X_l^T, l = 0..k-1  (one workgroup per l)
__global__ void __launch_bounds__(256) kern_ret_start_store(Pt p, const z2 *X, double *err) {
  __shared__ double red[256];
  const int n = p.n, l = blockIdx.x;
  BlkAcc acc;
  const int tid = threadIdx.x;
  const z2 *x = X + (size_t)l * n * n;
  acc.zero();
  blk_axpy(acc, n, x, 2, z2{1.0, 0.0});
  z2 *g = p.gr(p.T, p.T - l - 1);
  double d2 = blk_dist2(acc, n, g, red);
  blk_store(acc, n, g);
  if (tid == 0) atomicAdd(err, sqrt(d2));
}

// complete G^R(T,T-m) from its finished history sum (in acc)
template <typename SE>
__device__ void ret_finalize(const PtT<SE> &p, const WTab &w, BlkAcc &acc, BlkLds2<8> &s, int m, z2 *Qm,
                             const z2 *Minv, double *red, double *err) {
  const int n = p.n, T = p.T;
  blk_scale(acc, z2{p.dt, 0.0});
  for (int l = 1; l <= w.k + 1; l++) blk_axpy(acc, n, p.gr(T, T - m + l), 0, z2{0.0, -w.bd[l] / p.dt});
  __syncthreads();
  blk_store(acc, n, Qm);
  __syncthreads();
  BlkAcc x;
  x.zero();
  blk_mac2(x, s, n, Qm, 0, Minv + (size_t)(T - m) * n * n, 0, z2{1.0, 0.0});
  z2 *g = p.gr(T, T - m);
  double d2 = blk_dist2(x, n, g, red);
  blk_store(x, n, g);
  if (threadIdx.x == 0) atomicAdd(err, sqrt(d2));
}

// history sums from the start-up blocks l = 0..k, for m = k+1..T
// (QA: the dependency chain's own accumulator, zeroed here); three real
// products per complex one and two LDS stages, as the bulk updates
template <int NF, typename SE>
__device__ __forceinline__ void ret_qinit(BlkLds2<8> &s, double *red, const PtT<SE> &p, const WTab *__restrict__ wtab,
                                          z2 *Q, z2 *QA, const z2 *Minv, double *err) {
  const WTab &w = *wtab;
  const int m = w.k + 1 + blockIdx.x, T = p.T, n = NF ? NF : p.n;
  if (m > T) return;
  for (int e = threadIdx.x; e < n * n; e += 256) QA[(size_t)m * n * n + e] = z2{0.0, 0.0};
  BlkAcc3 a3;
  a3.zero();
  for (int l = 0; l <= w.k; l++)
    blk_mac2<NF>(a3, s, n, p.gr(T, T - l), 0, p.sr(T - l, T - m), 0, z2{greg(w, m, l), 0.0});
  BlkAcc acc = a3.value();
  z2 *Qm = Q + (size_t)m * n * n;
  if (m == w.k + 1) {
    ret_finalize(p, w, acc, s, m, Qm, Minv, red, err);
    return;
  }
  // the start blocks' share of the finishing terms of block m (added by the
  // chain links for every later block): -i/dt^2 bd_{m-j} G^R(T, T-j)
  for (int j = m - w.k - 1; j <= w.k; j++)
    blk_axpy(acc, n, p.gr(T, T - j), 0, z2{0.0, -w.bd[m - j] / (p.dt * p.dt)});
  blk_store(acc, n, Qm);
}

template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
kern_ret_qinit(PtT<SE> p, const WTab *__restrict__ wtab, z2 *Q, z2 *QA, const z2 *Minv, double *err) {
  __shared__ BlkLds2<8> s;
  __shared__ double red[256];
  if (p.n == BO_NMAX) ret_qinit<BO_NMAX>(s, red, p, wtab, Q, QA, Minv, err);
  else ret_qinit<0>(s, red, p, wtab, Q, QA, Minv, err);
}

// One link nn of the retarded recurrence: add block nn's contribution to
// the later blocks m of its panel and of the next one, and finish block
// nn+1.  Split by rows: workgroup (b, m) owns rows 16b..16b+15 of block m
// (its waves take the four column tiles).  G^R(T, T-m) depends only on the
// same rows of the history sum, so the bands never exchange data; the
// squared update of each finished band goes to d2[m * nb + b].
//
// The links are the dependency chain, so a launch's time is the latency of
// its finishing workgroups, and it runs beside the bulk panel updates of the
// side stream.  Hence: operands go straight into registers (no LDS staging);
// of the finishing terms -i/dt bd_l G^R(T, T-m+l) in Q' = dt Q - ..., each
// link adds those of block nn to the blocks up to k+1 ahead, so a finishing
// workgroup adds only the newest one; the products take three real ones per
// complex one (TileAcc3); and the waves raise their issue priority.
template <typename SE>
__device__ __forceinline__ void ret_step_band(const PtT<SE> &p, const WTab *__restrict__ wtab, int nn, z2 *QA,
                                              const z2 *QB, const z2 *Minv, double *d2, int b, int by, int nb) {
  __builtin_amdgcn_s_setprio(3);
  const WTab &w = *wtab;
  __shared__ z2 band[16][BO_NMAX + 1];
  __shared__ double red[4];
  // bands on x: the finishing workgroups (m = nn + 1) are the first
  // dispatched, ahead of the updates of later blocks
  const int m = nn + 1 + by, T = p.T, n = p.n;
  if (m > T) return;
  const int r0 = 16 * b, c0 = 16 * (threadIdx.x >> 6);
  const bool fin = m == nn + 1, bdt = m - nn <= w.k + 1;
  z2 *Qm = QA + (size_t)m * n * n;
  TileAcc acc, gt, qb, old;
  Frag fa;
  FragT<BO_NMAX / 4, SE> fb;
  tile_load(acc, n, Qm, r0, c0);
  frag_left(fa, n, p.gr(T, T - nn), 0, r0);
  frag_right(fb, n, p.sr(T - nn, T - m), 0, c0);
  if (bdt) tile_load(gt, n, p.gr(T, T - nn), r0, c0);
  if (fin) {
    tile_load(qb, n, QB + (size_t)m * n * n, r0, c0);
    tile_load(old, n, p.gr(T, T - m), r0, c0);
  }
  {
    TileAcc3 a3;
    a3.init(acc);
    frag_mac(a3, fa, fb, n, z2{greg(w, m, nn), 0.0});
    acc = a3.value();
  }
  if (bdt) {
    const z2 cl = {0.0, -w.bd[m - nn] / (p.dt * p.dt)};
#pragma unroll
    for (int q = 0; q < 4; q++) acc.set(q, zadd(acc.get(q), zmul(cl, gt.get(q))));
  }
  if (!fin) {
    tile_store(acc, n, Qm, r0, c0);
    return;
  }
  // finish the band: Q' = dt (QA + QB), X = Q' Minv_{T-m}
  Frag fm;
  frag_right(fm, n, Minv + (size_t)(T - m) * n * n, 0, c0);
  acc.re = (acc.re + qb.re) * p.dt;
  acc.im = (acc.im + qb.im) * p.dt;
  Frag fq;
  frag_left_band(fq, acc, band, n);
  TileAcc3 x3;
  x3.zero();
  frag_mac(x3, fq, fm, n, z2{1.0, 0.0});
  const TileAcc x = x3.value();
  double s2 = 0;
#pragma unroll
  for (int q = 0; q < 4; q++)
    if (tile_i(r0, q) < n && tile_j(c0) < n) {
      const z2 d = zsub(x.get(q), old.get(q));
      s2 += d.x * d.x + d.y * d.y;
    }
  const double dd = wg_sum(s2, red);
  tile_store(x, n, p.gr(T, T - m), r0, c0);
  if (threadIdx.x == 0) d2[(size_t)m * nb + b] = dd;
}

template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
kern_ret_step_band(PtT<SE> p, const WTab *__restrict__ wtab, int nn, z2 *QA, const z2 *QB, const z2 *Minv, double *d2) {
  ret_step_band(p, wtab, nn, QA, QB, Minv, d2, blockIdx.x, blockIdx.y, gridDim.x);
}

// the same link for every system of a group (system = blockIdx.z; bands
// beyond a system's own count exit)
template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
kern_ret_step_band_grp(const ChainArgT<SE> *args, const WTab *__restrict__ wtab, int nn) {
  const ChainArgT<SE> &a = args[blockIdx.z];
  if ((int)blockIdx.x >= a.nb) return;
  ret_step_band(a.p, wtab, nn, a.QA, a.Q, a.Minv, a.d2, blockIdx.x, blockIdx.y, a.nb);
}

// ---- chain links of one product: the history sums carried times M^-1
// A block is finished as X_m = dt S_m Minv_{T-m}, S_m its history sum.  Within
// a panel the chain carries R_m = dt S_m Minv_{T-m} instead: with the transfer
// blocks
//   B_{t,m} = dt (greg(m,t) Sigma^R(T-t, T-m) + [m-t <= k+1] c_{m-t}) Minv_{T-m}
// (c_l = -i bd_l / dt^2, the folded finishing terms) a link adds
// R_m += X_t B_{t,m}, and the finishing workgroup's X_{t+1} = R_{t+1} +
// X_t B_{t,t+1} is one product, where kern_ret_step_band takes a product, an
// exchange of the band through LDS and a second product.  The transfer blocks
// read Sigma^R rows before T and the cached inverses of times before T only,
// so they are formed once per time step (kern_ret_tblk) for every corrector
// pass.  At the start of a panel, kern_ret_xform turns the panel's sums into
// R (R_m = dt (Q_m + QA_m) Minv_{T-m}, in place in QA); the links add to the
// next panel's blocks in the plain form, as kern_ret_step_band does.
//
// Transfer block of link t to block m = t + d (d = 1..RP, m up to the end of
// t's panel): tb[t RP + d - 1]
template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) kern_ret_tblk(PtT<SE> p, const WTab *__restrict__ wtab, int RP,
                                                     const z2 *Minv, z2 *tb) {
  const WTab &w = *wtab;
  __shared__ BlkLds s;
  const int d = blockIdx.x + 1, t = w.k + 1 + blockIdx.y, m = t + d, T = p.T, n = p.n;
  const int q0 = w.k + 1 + RP * ((t - w.k - 1) / RP), q1 = min(q0 + RP, T);
  if (m > q1) return;
  const z2 *Mi = Minv + (size_t)(T - m) * n * n;
  BlkAcc3 acc;
  acc.zero();
  blk_mac(acc, s, n, p.sr(T - t, T - m), 0, Mi, 0, z2{p.dt * greg(w, m, t), 0.0});
  if (d <= w.k + 1) blk_axpy(acc, n, Mi, 0, z2{0.0, -w.bd[d] / p.dt});
  blk_store(acc.value(), n, tb + ((size_t)t * RP + d - 1) * n * n);
}

// R_m = dt (Q_m + QA_m) Minv_{T-m} for the blocks m of panel (q0, q1], in
// place in QA; workgroup (b, m) owns rows 16b..16b+15, its waves the column
// tiles
template <int NF>
__device__ __forceinline__ void ret_xform(const Pt &p, int q0, z2 *QA, const z2 *Q, const z2 *Minv) {
  const int b = blockIdx.x, m = q0 + 1 + blockIdx.y, T = p.T, n = NF ? NF : p.n;
  const int r0 = 16 * b, c0 = 16 * (threadIdx.x >> 6);
  const size_t es = (size_t)n * n;
  z2 *QAm = QA + (size_t)m * es;
  TileAcc3 x;
  x.zero();
#pragma unroll
  for (int h = 0; h < 2; h++) {
    FragT<BO_NMAX / 8> fa, fq;
    FragT<BO_NMAX / 8> fm;
    frag_left(fa, n, (const z2 *)QAm, 0, r0, h * BO_NMAX / 8);
    frag_left(fq, n, Q + (size_t)m * es, 0, r0, h * BO_NMAX / 8);
    frag_right(fm, n, Minv + (size_t)(T - m) * es, 0, c0, h * BO_NMAX / 8);
#pragma unroll
    for (int s = 0; s < BO_NMAX / 8; s++) fa.v[s] = zadd(fa.v[s], fq.v[s]);
    frag_mac(x, fa, fm, n, z2{p.dt, 0.0}, h * BO_NMAX / 8);
  }
  __syncthreads();  // every wave has read the band before any writes its tile
  tile_store(x.value(), n, QAm, r0, c0);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
kern_ret_xform(Pt p, int q0, z2 *QA, const z2 *Q, const z2 *Minv) {
  if (p.n == BO_NMAX) ret_xform<BO_NMAX>(p, q0, QA, Q, Minv);
  else ret_xform<0>(p, q0, QA, Q, Minv);
}

// link nn: blocks m of its panel (m <= q1) get R_m += X_nn B_{nn,m} and the
// finishing one (m = nn + 1) is X_m = R_m + X_nn B_{nn,m}; blocks of the next
// panel get the plain contribution of kern_ret_step_band
template <int NF, typename SE>
__device__ __forceinline__ void ret_link(const PtT<SE> &p, const WTab *__restrict__ wtab, int nn, int q1, int RP,
                                         z2 *QA, const z2 *tb, double *d2, double *red) {
  const WTab &w = *wtab;
  const int b = blockIdx.x, nb = gridDim.x, m = nn + 1 + blockIdx.y, T = p.T, n = NF ? NF : p.n;
  const int r0 = 16 * b, c0 = 16 * (threadIdx.x >> 6);
  const size_t es = (size_t)n * n;
  z2 *Qm = QA + (size_t)m * es;
  TileAcc3 acc;
  {
    TileAcc q;
    tile_load(q, n, Qm, r0, c0);
    acc.init(q);
  }
  Frag fa;
  frag_left(fa, n, p.gr(T, T - nn), 0, r0);
  if (m <= q1) {
    Frag fb;
    frag_right(fb, n, tb + ((size_t)nn * RP + (m - nn - 1)) * es, 0, c0);
    TileAcc old;
    const bool fin = m == nn + 1;
    if (fin) tile_load(old, n, p.gr(T, T - m), r0, c0);
    frag_mac(acc, fa, fb, n, z2{1.0, 0.0});
    const TileAcc x = acc.value();
    if (!fin) {
      tile_store(x, n, Qm, r0, c0);
      return;
    }
    double s2 = 0;
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (tile_i(r0, q) < n && tile_j(c0) < n) {
        const z2 dd = zsub(x.get(q), old.get(q));
        s2 += dd.x * dd.x + dd.y * dd.y;
      }
    const double dsum = wg_sum(s2, red);
    tile_store(x, n, p.gr(T, T - m), r0, c0);
    if (threadIdx.x == 0) d2[(size_t)m * nb + b] = dsum;
    return;
  }
  FragT<BO_NMAX / 4, SE> fb;
  frag_right(fb, n, p.sr(T - nn, T - m), 0, c0);
  frag_mac(acc, fa, fb, n, z2{greg(w, m, nn), 0.0});
  if (m - nn <= w.k + 1) {
    TileAcc gt;
    tile_load(gt, n, p.gr(T, T - nn), r0, c0);
    const z2 cl = {0.0, -w.bd[m - nn] / (p.dt * p.dt)};
#pragma unroll
    for (int q = 0; q < 4; q++) acc.add(q, zmul(cl, gt.get(q)));
  }
  tile_store(acc.value(), n, Qm, r0, c0);
}

template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
kern_ret_link(PtT<SE> p, const WTab *__restrict__ wtab, int nn, int q1, int RP, z2 *QA, const z2 *tb, double *d2) {
  __builtin_amdgcn_s_setprio(3);
  __shared__ double red[4];
  if (p.n == BO_NMAX) ret_link<BO_NMAX>(p, wtab, nn, q1, RP, QA, tb, d2, red);
  else ret_link<0>(p, wtab, nn, q1, RP, QA, tb, d2, red);
}

// ---- a whole panel of the dependency chain in one launch
// The links of one panel [q0, q1) as one launch instead of one launch each:
// workgroup (m, b) owns band b of block m for one m in (q0, q1x] and adds the
// contributions of the blocks nn in [q0, min(m, q1)) one after the other, as
// they are finished; a block m <= q1 is then finished by it, a later one (in
// the next panel) leaves its sum in the chain accumulator.  Block q0 comes
// from an earlier launch; every later one is handed over inside the launch:
// its finishing workgroup stores the band, makes its stores visible at device
// scope (L2 write-back, for workgroups on other XCDs) and sets its flag to
// this phase's epoch, and a reader polls that flag and then invalidates its
// caches before it reads the band.
//
// Workgroups take (m, b) from a ticket drawn when they start, in the order of
// m, so a flag is only ever awaited from a workgroup that started after the
// one that sets it: the wait ends.  As a guard against a fault elsewhere, a
// wait gives up after about 2^21 polls (some 100 ms), sets *fail to 1 and
// carries on without the block, and so does every later wait of the step
// (*fail is zeroed with the step's update norms); the host turns it into an
// error.  ctr[0] is the ticket counter and ctr[1] counts the
// workgroups that are done: the last one zeroes both for the next launch.
// The epoch is read from device memory (kern_chain_epoch bumps it once per
// recurrence), so that a replayed graph waits for this replay's blocks.
constexpr unsigned CHAIN_SPIN_MAX = 1u << 21;

__global__ void kern_chain_epoch(unsigned *epoch) { *epoch += 1u; }

__device__ __forceinline__ int chain_ticket(unsigned *ctr) {
  __shared__ int s_ticket;
  if (threadIdx.x == 0) s_ticket = (int)atomicAdd(ctr, 1u);
  __syncthreads();
  return s_ticket;
}

__device__ __forceinline__ void chain_wait(const unsigned *flag, unsigned epoch, double *fail) {
  if (threadIdx.x == 0) {
    unsigned spins = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
      // after one wait has given up, the others of the step do not wait
      if (++spins == CHAIN_SPIN_MAX || __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0.0) {
        __hip_atomic_store(fail, 1.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// every wave's stores of the band are complete and written back before the
// flag is set
__device__ __forceinline__ void chain_publish(unsigned *flag, unsigned epoch) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(flag, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void chain_leave(unsigned *ctr) {
  if (threadIdx.x == 0 && atomicAdd(ctr + 1, 1u) == gridDim.x - 1) {
    atomicExch(ctr, 0u);
    atomicExch(ctr + 1, 0u);
  }
}

// the retarded panel (rows 16b..16b+15 of block m; the waves take the four
// column tiles); the terms are those of kern_ret_step_band
template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
kern_ret_chain(PtT<SE> p, const WTab *__restrict__ wtab, int q0, int q1, int nb, z2 *QA, const z2 *QB,
               const z2 *Minv, double *d2, unsigned *flags, unsigned *ctr, const unsigned *epoch_p, double *fail) {
  __builtin_amdgcn_s_setprio(3);
  const WTab &w = *wtab;
  __shared__ z2 band[16][BO_NMAX + 1];
  __shared__ double red[4];
  const int t = chain_ticket(ctr);
  const unsigned epoch = *epoch_p;
  const int m = q0 + 1 + t / nb, b = t % nb, T = p.T, n = p.n;
  const int r0 = 16 * b, c0 = 16 * (threadIdx.x >> 6);
  const bool fin = m <= q1;
  z2 *Qm = QA + (size_t)m * n * n;
  TileAcc3 acc;
  {
    TileAcc q;
    tile_load(q, n, Qm, r0, c0);
    acc.init(q);
  }
  const int nn1 = min(m, q1);
  for (int nn = q0; nn < nn1; nn++) {
    FragT<BO_NMAX / 4, SE> fb;
    frag_right(fb, n, p.sr(T - nn, T - m), 0, c0);  // ahead of the wait: it does not depend on the chain
    if (nn > q0) chain_wait(flags + (size_t)nn * nb + b, epoch, fail);
    Frag fa;
    frag_left(fa, n, p.gr(T, T - nn), 0, r0);
    frag_mac(acc, fa, fb, n, z2{greg(w, m, nn), 0.0});
    if (m - nn <= w.k + 1) {
      TileAcc gt;
      tile_load(gt, n, p.gr(T, T - nn), r0, c0);
      const z2 cl = {0.0, -w.bd[m - nn] / (p.dt * p.dt)};
#pragma unroll
      for (int q = 0; q < 4; q++) acc.add(q, zmul(cl, gt.get(q)));
    }
  }
  if (!fin) {
    tile_store(acc.value(), n, Qm, r0, c0);
    chain_leave(ctr);
    return;
  }
  // finish the band: Q' = dt (QA + QB), X = Q' Minv_{T-m}
  Frag fm;
  TileAcc qb, old;
  frag_right(fm, n, Minv + (size_t)(T - m) * n * n, 0, c0);
  tile_load(qb, n, QB + (size_t)m * n * n, r0, c0);
  tile_load(old, n, p.gr(T, T - m), r0, c0);
  TileAcc a = acc.value();
  a.re = (a.re + qb.re) * p.dt;
  a.im = (a.im + qb.im) * p.dt;
  Frag fq;
  frag_left_band(fq, a, band, n);
  TileAcc3 x3;
  x3.zero();
  frag_mac(x3, fq, fm, n, z2{1.0, 0.0});
  const TileAcc x = x3.value();
  double s2 = 0;
#pragma unroll
  for (int q = 0; q < 4; q++)
    if (tile_i(r0, q) < n && tile_j(c0) < n) {
      const z2 d = zsub(x.get(q), old.get(q));
      s2 += d.x * d.x + d.y * d.y;
    }
  const double dd = wg_sum(s2, red);
  tile_store(x, n, p.gr(T, T - m), r0, c0);
  if (threadIdx.x == 0) d2[(size_t)m * nb + b] = dd;
  chain_publish(flags + (size_t)m * nb + b, epoch);
  chain_leave(ctr);
}

// err += sum over m in [m0, m1] of sqrt(sum_b d2[m][b])  (one workgroup)
__global__ void kern_band_err(const double *d2, int m0, int m1, int nb, double *err) {
  __shared__ double red[256];
  const int tid = threadIdx.x;
  double s = 0;
  for (int m = m0 + tid; m <= m1; m += 256) {
    double t = 0;
    for (int b = 0; b < nb; b++) t += d2[(size_t)m * nb + b];
    s += sqrt(t);
  }
  red[tid] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) *err += red[0];
}

// contributions of the finished blocks nn in [q0, q1) to every m in (q1, T]:
// the history sums stay in registers across the whole panel.  8-deep K
// slices in two LDS stages (blk_mac2), both operands fetched a slice ahead,
// three workgroups a CU -- at nao = 64 the Dyson step took 2.6 ms longer with
// 32-deep slices in one stage at two a CU, 1.1 ms longer with 16-deep ones in
// two stages; a group of systems of small nao ran 1% faster than with 16-deep
// slices in one stage and the shared G^R block read only when staged.
template <int NF, typename SE>
__device__ __forceinline__ void ret_panel(BlkLds2<8> &s, const PtT<SE> &p, const WTab *__restrict__ wtab,
                                          int q0, int q1, int m0, z2 *Q, int bx) {
  const WTab &w = *wtab;
  const int m = m0 + bx, T = p.T, n = p.n;
  if (m > T) return;
  z2 *Qm = Q + (size_t)m * n * n;
  BlkAcc3 acc;
  {
    BlkAcc q;
    blk_load(q, n, Qm);
    acc.init(q);
  }
  for (int nn = q0; nn < q1; nn++)
    blk_mac2<NF>(acc, s, n, p.gr(T, T - nn), 0, p.sr(T - nn, T - m), 0, z2{greg(w, m, nn), 0.0});
  __syncthreads();
  blk_store(acc.value(), n, Qm);
}

template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
kern_ret_panel(PtT<SE> p, const WTab *__restrict__ wtab, int q0, int q1, int m0, z2 *Q) {
  __shared__ BlkLds2<8> s;
  if (p.n == BO_NMAX) ret_panel<BO_NMAX>(s, p, wtab, q0, q1, m0, Q, blockIdx.x);
  else ret_panel<0>(s, p, wtab, q0, q1, m0, Q, blockIdx.x);
}

template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
kern_ret_panel_grp(const ChainArgT<SE> *args, const WTab *__restrict__ wtab, int q0, int q1, int m0) {
  const ChainArgT<SE> &a = args[blockIdx.y];
  __shared__ BlkLds2<8> s;
  if (a.p.n == BO_NMAX) ret_panel<BO_NMAX>(s, a.p, wtab, q0, q1, m0, a.Q, blockIdx.x);
  else ret_panel<0>(s, a.p, wtab, q0, q1, m0, a.Q, blockIdx.x);
}

// kern_ret_panel split by rows, for the late panels whose (T - q1) blocks
// alone leave most of the chip idle: workgroup (m, b) owns rows
// 16b..16b+15 of block m, its waves the four column tiles, four times the
// workgroups at a quarter of the work each
template <typename SE>
__device__ __forceinline__ void ret_panel_band(const PtT<SE> &p, const WTab *__restrict__ wtab, int q0, int q1, int m0,
                                               z2 *Q, int bx, int by) {
  const WTab &w = *wtab;
  const int m = m0 + bx, T = p.T, n = p.n;
  if (m > T) return;
  const int r0 = 16 * by, c0 = 16 * (threadIdx.x >> 6);
  z2 *Qm = Q + (size_t)m * n * n;
  TileAcc3 acc;
  {
    TileAcc q;
    tile_load(q, n, Qm, r0, c0);
    acc.init(q);
  }
  for (int nn = q0; nn < q1; nn++)
#pragma unroll
    for (int h = 0; h < 2; h++) {
      FragT<BO_NMAX / 8> fa;
      FragT<BO_NMAX / 8, SE> fb;
      frag_left(fa, n, p.gr(T, T - nn), 0, r0, h * BO_NMAX / 8);
      frag_right(fb, n, p.sr(T - nn, T - m), 0, c0, h * BO_NMAX / 8);
      frag_mac(acc, fa, fb, n, z2{greg(w, m, nn), 0.0}, h * BO_NMAX / 8);
    }
  tile_store(acc.value(), n, Qm, r0, c0);
}

template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
kern_ret_panel_band(PtT<SE> p, const WTab *__restrict__ wtab, int q0, int q1, int m0, z2 *Q) {
  ret_panel_band(p, wtab, q0, q1, m0, Q, blockIdx.x, blockIdx.y);
}

template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
kern_ret_panel_band_grp(const ChainArgT<SE> *args, const WTab *__restrict__ wtab, int q0, int q1, int m0) {
  const ChainArgT<SE> &a = args[blockIdx.z];
  if ((int)blockIdx.y >= a.nb) return;
  ret_panel_band(a.p, wtab, q0, q1, m0, a.Q, blockIdx.x, blockIdx.y);
}

// ------------------------------------------------------------ mixing (tv)
// src/dyson_step.h:130-169 with Ctv_tstp/CTV1/Conv.mixing
// (src/dyson_integrals.h:80-116, 254-324): partial sums over chunks of the
// time history, one workgroup per (node m, chunk s).
template <int NF, typename SE>
__device__ __forceinline__ void tv_rhs(BlkLds2<8> &s, const PtT<SE> &p, const WTab *__restrict__ wtab, const z2 *mixK,
                                       int nchunk, z2 *part) {
  const WTab &w = *wtab;
  const int m = blockIdx.x, ch = blockIdx.y, n = p.n, T = p.T, np = p.np;
  // terms t < T: the history j = t (G^rceil(T,.) is the unknown); t >= T:
  // the mixing node jp = t - T; chunks of equal length over both
  const int nterms = T + np;
  const int per = (nterms + nchunk - 1) / nchunk;
  const int t0 = ch * per, t1 = min(nterms, t0 + per);
  BlkAcc3 a3;
  a3.zero();
  for (int j = t0; j < min(t1, T); j++)
    blk_mac2<NF>(a3, s, n, p.sr(T, j), 0, p.gtv(j, m), 0, z2{p.dt * greg(w, T, j), 0.0});
  for (int jp = max(t0, T) - T; jp < t1 - T; jp++)
    blk_mac2<NF>(a3, s, n, p.stv(T, jp), 0, mixK + ((size_t)m * np + jp) * p.es(), 0, z2{1.0, 0.0});
  BlkAcc acc = a3.value();
  if (ch == 0)
    for (int l = 1; l <= w.k + 1; l++) blk_axpy(acc, n, p.gtv(T - l, m), 0, z2{0.0, -w.bd[l] / p.dt});
  blk_store(acc, n, part + ((size_t)ch * np + m) * p.es());
}

template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
kern_tv_rhs(PtT<SE> p, const WTab *__restrict__ wtab, const z2 *mixK, int nchunk, z2 *part) {
  __shared__ BlkLds2<8> s;
  if (p.n == BO_NMAX) tv_rhs<BO_NMAX>(s, p, wtab, mixK, nchunk, part);
  else tv_rhs<0>(s, p, wtab, mixK, nchunk, part);
}

template <typename SE>
__global__ void kern_tv_matrix(PtT<SE> p, const WTab *__restrict__ wtab, z2 *A, z2 *B) {
  const WTab &w = *wtab;
  const int n = p.n;
  const size_t es = p.es();
  for (int e = threadIdx.x; e < (int)es; e += blockDim.x) {
    int a = e / n, b = e % n;
    z2 hv = p.h(p.T)[e], sv = ldz(p.sr(p.T, p.T) + (e));
    double om0 = w.Omega[w.k * (w.k + 1)];
    z2 v = {-hv.x - p.dt * om0 * sv.x, -hv.y - p.dt * om0 * sv.y};
    if (a == b) { v.x += p.mu; v.y += w.bd[0] / p.dt; }
    A[e] = v;
    B[e] = (a == b) ? z2{1.0, 0.0} : z2{0.0, 0.0};
  }
}

__global__ void __launch_bounds__(256) kern_tv_fin(Pt p, const z2 *part, int nchunk, const z2 *Minv, z2 *scr,
                                                  double *err) {
  __shared__ BlkLds s;
  __shared__ double red[256];
  const int m = blockIdx.x, n = p.n, np = p.np;
  BlkAcc acc;
  acc.zero();
  for (int ch = 0; ch < nchunk; ch++) blk_axpy(acc, n, part + ((size_t)ch * np + m) * p.es(), 0, z2{1.0, 0.0});
  z2 *sc = scr + (size_t)m * p.es();
  blk_store(acc, n, sc);
  __syncthreads();
  BlkAcc x;
  x.zero();
  blk_mac(x, s, n, Minv, 0, sc, 0, z2{1.0, 0.0});
  z2 *g = p.gtv(p.T, m);
  double d2 = blk_dist2(x, n, g, red);
  blk_store(x, n, g);
  if (threadIdx.x == 0) atomicAdd(err, sqrt(d2));
}

// ------------------------------------------------------------ lesser
// Y_tau = beta/2 sum_tau' mass[tau][tau'] * sig i G^rceil(T, ntau-tau')^H
// (the right factor of Cles3_tstp, src/dyson_integrals.h:393-427)
// (workgroup (tau, chunk of 256 elements); thread f reads element f of each
// G^rceil block, so that the reads of a warp are contiguous, and writes the
// transposed element of Y once)
__global__ void kern_les_Y(Pt p, const double *mass, z2 *Y) {
  const int tau = blockIdx.x, n = p.n, np = p.np;
  const size_t es = p.es();
  const int f = blockIdx.y * blockDim.x + threadIdx.x;  // element (b, a) of G^rceil(T, .)
  if (f >= (int)es) return;
  const int b = f / n, a = f % n;
  z2 acc = {0.0, 0.0};
  for (int t2 = 0; t2 < np; t2++) {
    const double wgt = 0.5 * p.beta * mass[(size_t)tau * np + t2];
    const z2 g = p.gtv(p.T, np - 1 - t2)[f];  // element (a, b) of the adjoint
    // sig * i * conj(g)
    acc.x += wgt * (p.sig * g.y); acc.y += wgt * (p.sig * g.x);
  }
  Y[(size_t)tau * es + (size_t)a * n + b] = acc;
}

// Q_row = Cles3 + Cles2 (src/dyson_integrals.h:336-380, 393-427), partial
// sums over chunks of the (tau, j) terms.  8-deep K slices in two LDS stages
// (blk_mac2) keep both operands fetched a slice ahead at three workgroups a
// CU: 3% faster at nao = 64 than 16-deep slices in one stage with the shared
// G^R block read only when staged.
template <int NF, typename SE>
__device__ __forceinline__ void les_rhs(BlkLds2<8> &s, const PtT<SE> &p, const WTab *__restrict__ wtab, const z2 *Y,
                                        int nchunk, z2 *part) {
  const WTab &w = *wtab;
  const int row = blockIdx.x, ch = blockIdx.y, n = p.n, T = p.T, np = p.np;
  const int nterms = np + T + 1;
  const int per = (nterms + nchunk - 1) / nchunk;
  const int t0 = ch * per, t1 = min(nterms, t0 + per);
  BlkAcc3 acc;
  acc.zero();
  // three runs of terms, each with fixed operand forms (one inlined product
  // per run keeps the live registers to one copy):
  // tv terms t < np; then j = t - np: -Sigma^<(j,row)^H X_j for j < row and
  // Sigma^<(row,j) X_j for j >= row, with X_j = dt w(T,j) G^R(T,j)^H
  for (int t = t0; t < min(t1, np); t++)
    blk_mac2<NF>(acc, s, n, p.stv(row, t), 0, Y + (size_t)t * p.es(), 0, z2{1.0, 0.0});
  for (int t = max(t0, np); t < min(t1, np + row); t++) {
    const int j = t - np;
    blk_mac2<NF>(acc, s, n, p.sl(j, row), 1, p.gr(T, j), 1, z2{-p.dt * greg(w, T, j), 0.0});
  }
  for (int t = max(t0, np + row); t < t1; t++) {
    const int j = t - np;
    blk_mac2<NF>(acc, s, n, p.sl(row, j), 0, p.gr(T, j), 1, z2{p.dt * greg(w, T, j), 0.0});
  }
  blk_store(acc.value(), n, part + ((size_t)ch * (T + 1) + row) * p.es());
}

template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
kern_les_rhs(PtT<SE> p, const WTab *__restrict__ wtab, const z2 *Y, int nchunk, z2 *part) {
  __shared__ BlkLds2<8> s;
  if (p.n == BO_NMAX) les_rhs<BO_NMAX>(s, p, wtab, Y, nchunk, part);
  else les_rhs<0>(s, p, wtab, Y, nchunk, part);
}

__global__ void kern_sum_parts(const z2 *part, int nchunk, long long nel, z2 *out) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  z2 s = {0.0, 0.0};
  for (int c = 0; c < nchunk; c++) s = zadd(s, part[(size_t)c * nel + i]);
  out[i] = s;
}

// start-up of src/dyson_step.h:178-251: initial condition, kn x kn system.
// The matrix (times 0..k only) is built only when build_M is set, with the
// identity in MI for its inversion.
// G^<(0,T) = -G^rceil(T,0)^H ; X_0 = same   (one workgroup)
__global__ void __launch_bounds__(256) kern_les_x0(Pt p, z2 *X, double *err) {
  __shared__ double red[256];
  const int n = p.n, T = p.T;
  BlkAcc acc;
  acc.zero();
  blk_axpy(acc, n, p.gtv(T, 0), 1, z2{-1.0, 0.0});
  z2 *g = p.gl(0, T);
  double d2 = blk_dist2(acc, n, g, red);
  blk_store(acc, n, g);
  blk_store(acc, n, X);
  if (threadIdx.x == 0) atomicAdd(err, sqrt(d2));
}

// the start-up system (build_M) and its right-hand sides from X_0 (written
// by kern_les_x0), every element independent
template <typename SE>
__global__ void __launch_bounds__(256) kern_les_start(PtT<SE> p, const WTab *__restrict__ wtab, int build_M, const z2 *Q,
                                                      const z2 *X0, z2 *M, z2 *MI, z2 *Qs) {
  const WTab &w = *wtab;
  const int n = p.n, k = w.k, kn = k * n, k1 = k + 1;
  const long long gid = (long long)blockIdx.x * 256 + threadIdx.x, gsz = (long long)gridDim.x * 256;
  if (build_M)
    for (long long e = gid; e < (long long)kn * kn; e += gsz) {
      int row = (int)(e / kn), col = (int)(e % kn);
      int m = row / n + 1, a = row % n, l = col / n + 1, b = col % n;
      z2 v = {0.0, 0.0};
      if (a == b) v.y += w.pdiff[m * k1 + l] / p.dt;   // (i/dt) pd(m,l)
      if (m == l) {
        z2 hv = p.h(l)[a * n + b];
        v.x -= hv.x; v.y -= hv.y;
        if (a == b) v.x += p.mu;
      }
      double wt = p.dt * greg(w, m, l);
      if (m >= l) {
        z2 s = ldz(p.sr(m, l) + (a * n + b));
        v.x -= wt * s.x; v.y -= wt * s.y;
      } else {
        z2 s = ldz(p.sr(l, m) + (b * n + a));
        v.x += wt * s.x; v.y -= wt * s.y;
      }
      M[e] = v;
      MI[e] = (row == col) ? z2{1.0, 0.0} : z2{0.0, 0.0};
    }
  for (long long e = gid; e < (long long)kn * n; e += gsz) {
    int row = (int)(e / n), b = (int)(e % n);
    int m = row / n + 1, a = row % n;
    z2 v = Q[(size_t)m * n * n + a * n + b];
    double pd = w.pdiff[m * k1 + 0] / p.dt;
    z2 x0 = X0[a * n + b];
    v.x += pd * x0.y; v.y -= pd * x0.x;   // - i pd x0
    const SE *S = p.sr(m, 0);
    z2 sx = {0.0, 0.0};
    for (int c = 0; c < n; c++) sx = zadd(sx, zmul(ldz(S + a * n + c), X0[c * n + b]));
    double wt = p.dt * greg(w, m, 0);
    v.x += wt * sx.x; v.y += wt * sx.y;
    Qs[e] = v;
  }
}

template <typename SE>
__device__ void les_finalize(const PtT<SE> &p, const WTab &w, BlkAcc &acc, BlkLds2<8> &s, int m, z2 *Qm, z2 *X,
                             const z2 *Minv) {
  const int n = p.n;
  for (int l = 1; l <= w.k + 1; l++) blk_axpy(acc, n, X + (size_t)(m - l) * p.es(), 0, z2{0.0, -w.bd[l] / p.dt});
  __syncthreads();
  blk_store(acc, n, Qm);
  __syncthreads();
  BlkAcc x;
  x.zero();
  blk_mac2(x, s, n, Minv + (size_t)m * p.es(), 0, Qm, 0, z2{1.0, 0.0});
  blk_store(x, n, X + (size_t)m * p.es());
}

// (QA: the dependency chain's own accumulator, zeroed here); three real
// products per complex one and two LDS stages, as the bulk updates
template <int NF, typename SE>
__device__ __forceinline__ void les_qinit(BlkLds2<8> &s, const PtT<SE> &p, const WTab *__restrict__ wtab, z2 *Q,
                                          z2 *QA, z2 *X, const z2 *Minv) {
  const WTab &w = *wtab;
  const int m = w.k + 1 + blockIdx.x, T = p.T, n = NF ? NF : p.n;
  if (m > T) return;
  for (int e = threadIdx.x; e < n * n; e += 256) QA[(size_t)m * p.es() + e] = z2{0.0, 0.0};
  z2 *Qm = Q + (size_t)m * p.es();
  BlkAcc3 a3;
  {
    BlkAcc q;
    blk_load(q, n, Qm);
    a3.init(q);
  }
  for (int l = 0; l <= w.k; l++)
    blk_mac2<NF>(a3, s, n, p.sr(m, l), 0, X + (size_t)l * p.es(), 0, z2{p.dt * greg(w, m, l), 0.0});
  BlkAcc acc = a3.value();
  if (m == w.k + 1 && m <= T - 1) {
    les_finalize(p, w, acc, s, m, Qm, X, Minv);
    return;
  }
  // the start blocks' share of the finishing terms -i/dt bd_{m-j} X_j of
  // block m (not for m = T, which the diagonal step finishes)
  if (m <= T - 1)
    for (int j = m - w.k - 1; j <= w.k; j++) blk_axpy(acc, n, X + (size_t)j * p.es(), 0, z2{0.0, -w.bd[m - j] / p.dt});
  blk_store(acc, n, Qm);
}

template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
kern_les_qinit(PtT<SE> p, const WTab *__restrict__ wtab, z2 *Q, z2 *QA, z2 *X, const z2 *Minv) {
  __shared__ BlkLds2<8> s;
  if (p.n == BO_NMAX) les_qinit<BO_NMAX>(s, p, wtab, Q, QA, X, Minv);
  else les_qinit<0>(s, p, wtab, Q, QA, X, Minv);
}

// One link l of the lesser recurrence, split by columns: workgroup (b, m)
// owns columns 16b..16b+15 of block m (its waves take the four row tiles);
// X_m = Minv_m Q'_m keeps the columns of Q'_m apart.  Built as
// kern_ret_step_band (the finishing terms -i/dt bd_l X_{m-l} added ahead by
// link m-l).
template <typename SE>
__device__ __forceinline__ void les_step_band(const PtT<SE> &p, const WTab *__restrict__ wtab, int l, z2 *QA,
                                              const z2 *QB, z2 *X, const z2 *Minv, int b, int by) {
  __builtin_amdgcn_s_setprio(3);
  const WTab &w = *wtab;
  __shared__ z2 band[BO_NMAX][17];
  const int m = l + 1 + by, T = p.T, n = p.n;  // finishing bands first
  if (m > T) return;
  const int r0 = 16 * (threadIdx.x >> 6), c0 = 16 * b;
  const bool fin = m == l + 1 && m <= T - 1, bdt = m - l <= w.k + 1 && m <= T - 1;
  const size_t es = p.es();
  z2 *Qm = QA + (size_t)m * es;
  TileAcc acc, gt, qb;
  FragT<BO_NMAX / 4, SE> fa;
  Frag fb;
  tile_load(acc, n, Qm, r0, c0);
  frag_left(fa, n, p.sr(m, l), 0, r0);
  frag_right(fb, n, X + (size_t)l * es, 0, c0);
  if (bdt) tile_load(gt, n, X + (size_t)l * es, r0, c0);
  if (fin) tile_load(qb, n, QB + (size_t)m * es, r0, c0);
  {
    TileAcc3 a3;
    a3.init(acc);
    frag_mac(a3, fa, fb, n, z2{p.dt * greg(w, m, l), 0.0});
    acc = a3.value();
  }
  if (bdt) {
    const z2 cl = {0.0, -w.bd[m - l] / p.dt};
#pragma unroll
    for (int q = 0; q < 4; q++) acc.set(q, zadd(acc.get(q), zmul(cl, gt.get(q))));
  }
  if (!fin) {
    tile_store(acc, n, Qm, r0, c0);
    return;
  }
  Frag fm;
  frag_left(fm, n, Minv + (size_t)m * es, 0, r0);
  acc.re += qb.re;
  acc.im += qb.im;
  Frag fq;
  frag_right_band(fq, acc, band, n);
  TileAcc3 x3;
  x3.zero();
  frag_mac(x3, fm, fq, n, z2{1.0, 0.0});
  tile_store(x3.value(), n, X + (size_t)m * es, r0, c0);
}

template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
kern_les_step_band(PtT<SE> p, const WTab *__restrict__ wtab, int l, z2 *QA, const z2 *QB, z2 *X, const z2 *Minv) {
  les_step_band(p, wtab, l, QA, QB, X, Minv, blockIdx.x, blockIdx.y);
}

template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
kern_les_step_band_grp(const ChainArgT<SE> *args, const WTab *__restrict__ wtab, int l) {
  const ChainArgT<SE> &a = args[blockIdx.z];
  if ((int)blockIdx.x >= a.nb) return;
  les_step_band(a.p, wtab, l, a.QA, a.Q, a.X, a.Minv, blockIdx.x, blockIdx.y);
}

// ---- lesser chain links of one product, as kern_ret_link: X_m =
// Minv_m S_m, so the panel carries R_m = Minv_m S_m, with the transfer blocks
//   B_{m,l} = Minv_m (dt greg(m,l) Sigma^R(m,l) + [m-l <= k+1] c_{m-l}),
// c_l = -i bd_l / dt, multiplying X_l from the left (R_m += B_{m,l} X_l).
// Block T is finished by the diagonal step from the plain sums, so it always
// takes the plain contribution.  Transfer block of link l to m = l + d:
// tb[l RP + d - 1].
template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) kern_les_tblk(PtT<SE> p, const WTab *__restrict__ wtab, int RP,
                                                     const z2 *Minv, z2 *tb) {
  const WTab &w = *wtab;
  __shared__ BlkLds s;
  const int d = blockIdx.x + 1, l = w.k + 1 + blockIdx.y, m = l + d, T = p.T, n = p.n;
  const int q0 = w.k + 1 + RP * ((l - w.k - 1) / RP), q1 = min(q0 + RP, T);
  if (m > q1 || m > T - 1) return;
  const z2 *Mi = Minv + (size_t)m * n * n;
  BlkAcc3 acc;
  acc.zero();
  blk_mac(acc, s, n, Mi, 0, p.sr(m, l), 0, z2{p.dt * greg(w, m, l), 0.0});
  if (d <= w.k + 1) blk_axpy(acc, n, Mi, 0, z2{0.0, -w.bd[d] / p.dt});
  blk_store(acc.value(), n, tb + ((size_t)l * RP + d - 1) * n * n);
}

// R_m = Minv_m (Q_m + QA_m) for the blocks m of panel (q0, q1] (not T), in
// place in QA; workgroup (b, m) owns columns 16b..16b+15, its waves the row
// tiles
template <int NF>
__device__ __forceinline__ void les_xform(const Pt &p, int q0, z2 *QA, const z2 *Q, const z2 *Minv) {
  const int b = blockIdx.x, m = q0 + 1 + blockIdx.y, n = NF ? NF : p.n;
  const int r0 = 16 * (threadIdx.x >> 6), c0 = 16 * b;
  const size_t es = (size_t)n * n;
  z2 *QAm = QA + (size_t)m * es;
  TileAcc3 x;
  x.zero();
#pragma unroll
  for (int h = 0; h < 2; h++) {
    FragT<BO_NMAX / 8> fm, fa, fq;
    frag_left(fm, n, Minv + (size_t)m * es, 0, r0, h * BO_NMAX / 8);
    frag_right(fa, n, (const z2 *)QAm, 0, c0, h * BO_NMAX / 8);
    frag_right(fq, n, Q + (size_t)m * es, 0, c0, h * BO_NMAX / 8);
#pragma unroll
    for (int s = 0; s < BO_NMAX / 8; s++) fa.v[s] = zadd(fa.v[s], fq.v[s]);
    frag_mac(x, fm, fa, n, z2{1.0, 0.0}, h * BO_NMAX / 8);
  }
  __syncthreads();  // every wave has read the band before any writes its tile
  tile_store(x.value(), n, QAm, r0, c0);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
kern_les_xform(Pt p, int q0, z2 *QA, const z2 *Q, const z2 *Minv) {
  if (p.n == BO_NMAX) les_xform<BO_NMAX>(p, q0, QA, Q, Minv);
  else les_xform<0>(p, q0, QA, Q, Minv);
}

template <int NF, typename SE>
__device__ __forceinline__ void les_link(const PtT<SE> &p, const WTab *__restrict__ wtab, int l, int q1, int RP,
                                         z2 *QA, z2 *X, const z2 *tb) {
  const WTab &w = *wtab;
  const int b = blockIdx.x, m = l + 1 + blockIdx.y, T = p.T, n = NF ? NF : p.n;
  const int r0 = 16 * (threadIdx.x >> 6), c0 = 16 * b;
  const size_t es = (size_t)n * n;
  z2 *Qm = QA + (size_t)m * es;
  TileAcc3 acc;
  {
    TileAcc q;
    tile_load(q, n, Qm, r0, c0);
    acc.init(q);
  }
  Frag fb;
  frag_right(fb, n, X + (size_t)l * es, 0, c0);
  if (m <= q1 && m <= T - 1) {
    Frag fa;
    frag_left(fa, n, tb + ((size_t)l * RP + (m - l - 1)) * es, 0, r0);
    frag_mac(acc, fa, fb, n, z2{1.0, 0.0});
    tile_store(acc.value(), n, m == l + 1 ? X + (size_t)m * es : Qm, r0, c0);
    return;
  }
  FragT<BO_NMAX / 4, SE> fa;
  frag_left(fa, n, p.sr(m, l), 0, r0);
  frag_mac(acc, fa, fb, n, z2{p.dt * greg(w, m, l), 0.0});
  if (m - l <= w.k + 1 && m <= T - 1) {
    TileAcc gt;
    tile_load(gt, n, X + (size_t)l * es, r0, c0);
    const z2 cl = {0.0, -w.bd[m - l] / p.dt};
#pragma unroll
    for (int q = 0; q < 4; q++) acc.add(q, zmul(cl, gt.get(q)));
  }
  tile_store(acc.value(), n, Qm, r0, c0);
}

template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
kern_les_link(PtT<SE> p, const WTab *__restrict__ wtab, int l, int q1, int RP, z2 *QA, z2 *X, const z2 *tb) {
  __builtin_amdgcn_s_setprio(3);
  if (p.n == BO_NMAX) les_link<BO_NMAX>(p, wtab, l, q1, RP, QA, X, tb);
  else les_link<0>(p, wtab, l, q1, RP, QA, X, tb);
}

// the lesser panel in one launch, as kern_ret_chain (columns 16b..16b+15 of
// block m; the waves take the four row tiles; the terms of
// kern_les_step_band; block T is left to the diagonal step)
template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
kern_les_chain(PtT<SE> p, const WTab *__restrict__ wtab, int q0, int q1, int nb, z2 *QA, const z2 *QB, z2 *X,
               const z2 *Minv, unsigned *flags, unsigned *ctr, const unsigned *epoch_p, double *fail) {
  __builtin_amdgcn_s_setprio(3);
  const WTab &w = *wtab;
  __shared__ z2 band[BO_NMAX][17];
  const int t = chain_ticket(ctr);
  const unsigned epoch = *epoch_p;
  const int m = q0 + 1 + t / nb, b = t % nb, T = p.T, n = p.n;
  const int r0 = 16 * (threadIdx.x >> 6), c0 = 16 * b;
  const bool fin = m <= q1 && m <= T - 1;
  const size_t es = p.es();
  z2 *Qm = QA + (size_t)m * es;
  TileAcc3 acc;
  {
    TileAcc q;
    tile_load(q, n, Qm, r0, c0);
    acc.init(q);
  }
  const int l1 = min(m, q1);
  for (int l = q0; l < l1; l++) {
    FragT<BO_NMAX / 4, SE> fa;
    frag_left(fa, n, p.sr(m, l), 0, r0);  // ahead of the wait
    if (l > q0) chain_wait(flags + (size_t)l * nb + b, epoch, fail);
    Frag fb;
    frag_right(fb, n, X + (size_t)l * es, 0, c0);
    frag_mac(acc, fa, fb, n, z2{p.dt * greg(w, m, l), 0.0});
    if (m - l <= w.k + 1 && m <= T - 1) {
      TileAcc gt;
      tile_load(gt, n, X + (size_t)l * es, r0, c0);
      const z2 cl = {0.0, -w.bd[m - l] / p.dt};
#pragma unroll
      for (int q = 0; q < 4; q++) acc.add(q, zmul(cl, gt.get(q)));
    }
  }
  if (!fin) {
    tile_store(acc.value(), n, Qm, r0, c0);
    chain_leave(ctr);
    return;
  }
  Frag fm;
  TileAcc qb;
  frag_left(fm, n, Minv + (size_t)m * es, 0, r0);
  tile_load(qb, n, QB + (size_t)m * es, r0, c0);
  TileAcc a = acc.value();
  a.re += qb.re;
  a.im += qb.im;
  Frag fq;
  frag_right_band(fq, a, band, n);
  TileAcc3 x3;
  x3.zero();
  frag_mac(x3, fm, fq, n, z2{1.0, 0.0});
  tile_store(x3.value(), n, X + (size_t)m * es, r0, c0);
  chain_publish(flags + (size_t)m * nb + b, epoch);
  chain_leave(ctr);
}

// contributions of the finished X_l, l in [q0, q1), to every m in (q1, T]
// (LDS stages and slices chosen as in ret_panel)
template <int NF, typename SE>
__device__ __forceinline__ void les_panel(BlkLds2<8> &s, const PtT<SE> &p, const WTab *__restrict__ wtab,
                                          int q0, int q1, int m0, z2 *Q, const z2 *X, int bx) {
  const WTab &w = *wtab;
  const int m = m0 + bx, T = p.T, n = p.n;
  if (m > T) return;
  z2 *Qm = Q + (size_t)m * p.es();
  BlkAcc3 acc;
  {
    BlkAcc q;
    blk_load(q, n, Qm);
    acc.init(q);
  }
  for (int l = q0; l < q1; l++)
    blk_mac2<NF>(acc, s, n, p.sr(m, l), 0, X + (size_t)l * p.es(), 0, z2{p.dt * greg(w, m, l), 0.0});
  __syncthreads();
  blk_store(acc.value(), n, Qm);
}

template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
kern_les_panel(PtT<SE> p, const WTab *__restrict__ wtab, int q0, int q1, int m0, z2 *Q, const z2 *X) {
  __shared__ BlkLds2<8> s;
  if (p.n == BO_NMAX) les_panel<BO_NMAX>(s, p, wtab, q0, q1, m0, Q, X, blockIdx.x);
  else les_panel<0>(s, p, wtab, q0, q1, m0, Q, X, blockIdx.x);
}

template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
kern_les_panel_grp(const ChainArgT<SE> *args, const WTab *__restrict__ wtab, int q0, int q1, int m0) {
  const ChainArgT<SE> &a = args[blockIdx.y];
  __shared__ BlkLds2<8> s;
  if (a.p.n == BO_NMAX) les_panel<BO_NMAX>(s, a.p, wtab, q0, q1, m0, a.Q, a.X, blockIdx.x);
  else les_panel<0>(s, a.p, wtab, q0, q1, m0, a.Q, a.X, blockIdx.x);
}

// kern_les_panel split by columns (workgroup (m, b) owns columns
// 16b..16b+15 of block m), for the late panels, as kern_ret_panel_band
template <typename SE>
__device__ __forceinline__ void les_panel_band(const PtT<SE> &p, const WTab *__restrict__ wtab, int q0, int q1, int m0,
                                               z2 *Q, const z2 *X, int bx, int by) {
  const WTab &w = *wtab;
  const int m = m0 + bx, T = p.T, n = p.n;
  if (m > T) return;
  const int r0 = 16 * (threadIdx.x >> 6), c0 = 16 * by;
  const size_t es = p.es();
  z2 *Qm = Q + (size_t)m * es;
  TileAcc3 acc;
  {
    TileAcc q;
    tile_load(q, n, Qm, r0, c0);
    acc.init(q);
  }
  for (int l = q0; l < q1; l++)
#pragma unroll
    for (int h = 0; h < 2; h++) {
      FragT<BO_NMAX / 8, SE> fa;
      FragT<BO_NMAX / 8> fb;
      frag_left(fa, n, p.sr(m, l), 0, r0, h * BO_NMAX / 8);
      frag_right(fb, n, X + (size_t)l * es, 0, c0, h * BO_NMAX / 8);
      frag_mac(acc, fa, fb, n, z2{p.dt * greg(w, m, l), 0.0}, h * BO_NMAX / 8);
    }
  tile_store(acc.value(), n, Qm, r0, c0);
}

template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
kern_les_panel_band(PtT<SE> p, const WTab *__restrict__ wtab, int q0, int q1, int m0, z2 *Q, const z2 *X) {
  les_panel_band(p, wtab, q0, q1, m0, Q, X, blockIdx.x, blockIdx.y);
}

template <typename SE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
kern_les_panel_band_grp(const ChainArgT<SE> *args, const WTab *__restrict__ wtab, int q0, int q1, int m0) {
  const ChainArgT<SE> &a = args[blockIdx.z];
  if ((int)blockIdx.y >= a.nb) return;
  les_panel_band(a.p, wtab, q0, q1, m0, a.Q, a.X, blockIdx.x, blockIdx.y);
}

// diagonal step, src/dyson_step.h:283-310
template <typename SE>
__global__ void __launch_bounds__(256) kern_les_diag(PtT<SE> p, const WTab *__restrict__ wtab, z2 *Q, const z2 *QA, z2 *X,
                                                     z2 *scr) {
  const WTab &w = *wtab;
  __shared__ BlkLds s;
  const int T = p.T, n = p.n;
  const size_t es = p.es();
  z2 *XT = X + (size_t)T * es;
  BlkAcc acc;
  acc.zero();
  for (int jj = 0; jj <= w.k; jj++) blk_axpy(acc, n, p.gl(T - jj - 1, T - jj - 1), 0, z2{w.ex[jj], 0.0});
  blk_store(acc, n, XT);
  __syncthreads();
  // Q_T += dt w(T,T) Sigma^R(T,T) X_T + (hmf(T) - mu) X_T
  blk_load(acc, n, Q + (size_t)T * es);
  blk_axpy(acc, n, QA + (size_t)T * es, 0, z2{1.0, 0.0});
  blk_mac(acc, s, n, p.sr(T, T), 0, XT, 0, z2{p.dt * greg(w, T, T), 0.0});
  blk_mac(acc, s, n, p.h(T), 0, XT, 0, z2{1.0, 0.0});
  blk_axpy(acc, n, XT, 0, z2{-p.mu, 0.0});
  __syncthreads();
  blk_store(acc, n, scr);
  __syncthreads();
  // Q0 = -i (Q + Q^H) - 1/dt sum_l bd_l G^<(T-l,T-l);  X_T = Q0 dt / bd_0
  BlkAcc q0;
  q0.zero();
  blk_axpy(q0, n, scr, 0, z2{0.0, -1.0});
  blk_axpy(q0, n, scr, 1, z2{0.0, -1.0});
  for (int l = 1; l <= w.k + 1; l++) blk_axpy(q0, n, p.gl(T - l, T - l), 0, z2{-w.bd[l] / p.dt, 0.0});
  blk_scale(q0, z2{p.dt / w.bd[0], 0.0});
  __syncthreads();
  blk_store(q0, n, XT);
}

__global__ void __launch_bounds__(256) kern_les_store(Pt p, const z2 *X, double *err) {
  __shared__ double red[256];
  const int l = blockIdx.x, n = p.n;
  BlkAcc acc;
  blk_load(acc, n, X + (size_t)l * p.es());
  z2 *g = p.gl(l, p.T);
  double d2 = blk_dist2(acc, n, g, red);
  blk_store(acc, n, g);
  if (threadIdx.x == 0) atomicAdd(err, sqrt(d2));
}

// ------------------------------------------------------------ predictor
// dyson::Extrapolate (src/dyson.cpp:47-117); one thread per output element
__global__ void kern_extrap(Pt p, const WTab *__restrict__ wtab, int which) {
  const WTab &w = *wtab;
  const int n = p.n, k = w.k, T = p.T, np = p.np;
  const size_t es = p.es();
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nblk = which == 0 ? np : T + 1;
  if (idx >= (long long)nblk * es) return;
  const int l = (int)(idx / es), e = (int)(idx % es), a = e / n, b = e % n;
  const double *ex = w.ex;
  z2 r = {0.0, 0.0};
  if (which == 0) {  // tv
    for (int j = 0; j <= k; j++) r = zadd(r, zscale(p.gtv(T - j - 1, l)[e], ex[j]));
    p.gtv(T, l)[e] = r;
  } else if (which == 1) {  // ret column index l
    if (l > k) {  // l = T - l' with l' in [0, T-k-1]
      const int lp = T - l;
      for (int j = 0; j <= k; j++) r = zadd(r, zscale(p.gr(T - j - 1, T - lp - j - 1)[e], ex[j]));
    } else {
      const int jcut = (l <= T - k - 1) ? k : (T - l - 1);
      for (int j = 0; j <= jcut; j++) r = zadd(r, zscale(p.gr(T - j - 1, l)[e], ex[j]));
      for (int j = jcut + 1; j <= k; j++) r = zsub(r, zscale(zconj(p.gr(l, T - j - 1)[b * n + a]), ex[j]));
    }
    p.gr(T, l)[e] = r;
  } else {  // les row index l
    if (l == 0) {
      r = zscale(zconj(p.gtv(T, 0)[b * n + a]), -1.0);
    } else if (l <= k) {
      const int jcut = (k > T - l - 1) ? (T - l - 1) : k;
      for (int j = 0; j <= jcut; j++) r = zadd(r, zscale(p.gl(l, T - 1 - j)[e], ex[j]));
      for (int j = jcut + 1; j <= k; j++) r = zsub(r, zscale(zconj(p.gl(T - 1 - j, l)[b * n + a]), ex[j]));
    } else {
      for (int j = 0; j <= k; j++) r = zadd(r, zscale(p.gl(l - j - 1, T - j - 1)[e], ex[j]));
    }
    p.gl(l, T)[e] = r;
  }
}

// ------------------------------------------------------------ HF
// X_Q = sum_kl rho_kl B[Q][k][l]; rho = i sig G^<(T,T)
__global__ void kern_hf_x(Pt p, const double *Bq, z2 *Xq) {
  __shared__ double rx[256], ry[256];
  const int Q = blockIdx.x;
  const size_t es = p.es();
  const z2 *gl = p.gl(p.T, p.T);
  double sx = 0, sy = 0;
  for (int e = threadIdx.x; e < (int)es; e += 256) {
    z2 g = gl[e];
    z2 rho = {-p.sig * g.y, p.sig * g.x};
    double v = Bq[(size_t)Q * es + e];
    sx += rho.x * v; sy += rho.y * v;
  }
  rx[threadIdx.x] = sx; ry[threadIdx.x] = sy;
  __syncthreads();
  for (int w2 = 128; w2 > 0; w2 >>= 1) {
    if (threadIdx.x < w2) { rx[threadIdx.x] += rx[threadIdx.x + w2]; ry[threadIdx.x] += ry[threadIdx.x + w2]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) Xq[Q] = {rx[0], ry[0]};
}

// hmf(T)_ij = h0_ij + 2 sum_Q X_Q B[Q][i][j] - sum_(l,k) rho_kl U'[(l,k)][(i,j)]:
// a matrix-vector product over naux + n^2 terms, HBM-bound on U' (n^4
// doubles).  The terms are split over gridDim.y workgroup rows (partial sums
// part[s][e]) so that the whole chip streams U', then kern_hf_sum adds them.
__global__ void kern_hf_h(Pt p, const double *Bq, const double *U, const z2 *Xq, int naux, z2 *part) {
  const int n = p.n, S = gridDim.y, s = blockIdx.y;
  const size_t es = p.es();
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)es) return;
  const int terms = naux + n * n, per = (terms + S - 1) / S;
  const int t0 = s * per, t1 = min(terms, t0 + per);
  z2 v = {0.0, 0.0};
  for (int Q = t0; Q < min(t1, naux); Q++) {
    const double b = Bq[(size_t)Q * es + e];
    v.x += 2 * Xq[Q].x * b; v.y += 2 * Xq[Q].y * b;
  }
  const z2 *gl = p.gl(p.T, p.T);
  for (int t = max(t0, naux); t < t1; t++) {
    const int lk = t - naux, l = lk / n, k = lk % n;
    const z2 g = gl[k * n + l];
    const z2 rho = {-p.sig * g.y, p.sig * g.x};
    const double u = U[(size_t)lk * es + e];
    v.x -= rho.x * u; v.y -= rho.y * u;
  }
  part[(size_t)s * es + e] = v;
}

__global__ void kern_hf_sum(Pt p, const double *h0, const z2 *part, int S, z2 *hmf) {
  const size_t es = p.es();
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)es) return;
  z2 v = {h0[e], 0.0};
  for (int s = 0; s < S; s++) v = zadd(v, part[(size_t)s * es + e]);
  hmf[(size_t)p.T * es + e] = v;
}

// ------------------------------------------------------------ energy
// partial traces of energy_conv's history terms (src/dyson_integrals.h:37-58)
template <typename SE>
__global__ void kern_energy(PtT<SE> p, const WTab *__restrict__ wtab, int top, const double *mass, double *out) {
  const WTab &w = *wtab;
  __shared__ double rx[256], ry[256];
  const int n = p.n, T = p.T, np = p.np;
  const size_t es = p.es();
  double sx = 0, sy = 0;
  const int i = blockIdx.x;
  auto trT = [&](const SE *S, const z2 *G, int e) {  // S_ab G_ba
    int a = e / n, b = e % n;
    return zmul(ldz(S + e), G[b * n + a]);
  };
  if (i <= top) {
    double wt = greg(w, T, i);
    for (int e = threadIdx.x; e < (int)es; e += 256) {
      z2 t1, t2;
      if (i <= T) {
        t1 = zscale(trT(p.sr(T, i), p.gl(i, T), e), wt);
        t2 = zscale(zconj(trT(p.sl(i, T), p.gr(T, i), e)), -wt);
      } else {
        t1 = zscale(zconj(trT(p.sr(i, T), p.gl(T, i), e)), wt);
        t2 = zscale(trT(p.sl(T, i), p.gr(i, T), e), -wt);
      }
      sx += (t1.x + t2.x) * p.dt; sy += (t1.y + t2.y) * p.dt;
    }
  } else {
    // Conv.energy row: tau index jt = i - top - 1
    int jt = i - top - 1;
    for (int e = threadIdx.x; e < (int)es; e += 256) {
      z2 S = p.stv(T, jt)[e];
      for (int kk = 0; kk < np; kk++) {
        double wq = 0.5 * p.beta * mass[(size_t)jt * np + (np - 1 - kk)];
        z2 g = zconj(p.gtv(T, kk)[e]);
        z2 v = zmul(S, g);
        // res2 = -sig * (...) then times -i
        double f = -p.sig * wq;
        sx += f * v.y; sy += -f * v.x;
      }
    }
  }
  rx[threadIdx.x] = sx; ry[threadIdx.x] = sy;
  __syncthreads();
  for (int w2 = 128; w2 > 0; w2 >>= 1) {
    if (threadIdx.x < w2) { rx[threadIdx.x] += rx[threadIdx.x + w2]; ry[threadIdx.x] += ry[threadIdx.x + w2]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out[2 * i] = rx[0]; out[2 * i + 1] = ry[0]; }
}

// ------------------------------------------------------------ bootstrap
// src/dyson_boot.h:93-179: boundary condition, then the kn x kn system (the
// same for every tau node) with one right-hand side per node.
__global__ void kern_boot_tv0(Pt p, double *err) {
  __shared__ double red[256];
  const int m = blockIdx.x, n = p.n;
  BlkAcc acc;
  acc.zero();
  blk_axpy(acc, n, p.Gmat + (size_t)(p.np - 1 - m) * p.es(), 0, z2{0.0, (double)p.sig});
  z2 *g = p.gtv(0, m);
  double d2 = blk_dist2(acc, n, g, red);
  blk_store(acc, n, g);
  if (threadIdx.x == 0) atomicAdd(err, sqrt(d2));
}

template <typename SE>
__global__ void kern_boot_matrix(PtT<SE> p, const WTab *__restrict__ wtab, z2 *M, z2 *B) {
  const WTab &w = *wtab;
  const int n = p.n, k = w.k, kn = k * n, k1 = k + 1;
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)kn * kn) return;
  int row = (int)(e / kn), col = (int)(e % kn);
  int nn = row / n + 1, a = row % n, l = col / n + 1, b = col % n;
  z2 v = {0.0, 0.0};
  if (a == b) v.y += w.pdiff[nn * k1 + l] / p.dt;
  if (l == nn) {
    z2 hv = p.h(l)[a * n + b];
    v.x -= hv.x; v.y -= hv.y;
    if (a == b) v.x += p.mu;
  }
  double wt = p.dt * greg(w, nn, l);
  if (nn >= l) {
    z2 s = ldz(p.sr(nn, l) + (a * n + b));
    v.x -= wt * s.x; v.y -= wt * s.y;
  } else {
    z2 s = ldz(p.sr(l, nn) + (b * n + a));
    v.x += wt * s.x; v.y -= wt * s.y;
  }
  M[e] = v;
  B[e] = (row == col) ? z2{1.0, 0.0} : z2{0.0, 0.0};
}

// right-hand side of node m, step nn: tv_it_conv + derivative + integral terms
template <typename SE>
__global__ void __launch_bounds__(256) kern_boot_rhs(PtT<SE> p, const WTab *__restrict__ wtab, const z2 *mixK, z2 *R) {
  const WTab &w = *wtab;
  __shared__ BlkLds s;
  const int m = blockIdx.x, nn = blockIdx.y + 1, n = p.n, np = p.np, k1 = w.k + 1;
  BlkAcc acc;
  acc.zero();
  for (int i = 0; i < np; i++) blk_mac(acc, s, n, p.stv(nn, i), 0, mixK + ((size_t)m * np + i) * p.es(), 0, z2{1.0, 0.0});
  const z2 *g0 = p.gtv(0, m);
  blk_axpy(acc, n, g0, 0, z2{0.0, -w.pdiff[nn * k1] / p.dt});
  blk_mac(acc, s, n, p.sr(nn, 0), 0, g0, 0, z2{p.dt * greg(w, nn, 0), 0.0});
  blk_store(acc, n, R + ((size_t)(nn - 1) * np + m) * p.es());
}

// G^rceil(l+1, m) = sum_n' Minv[l][n'] R[n'][m]
__global__ void __launch_bounds__(256) kern_boot_apply(Pt p, int k, const z2 *Minv, const z2 *R, double *err) {
  __shared__ BlkLds s;
  __shared__ double red[256];
  const int m = blockIdx.x, l = blockIdx.y, n = p.n, np = p.np, kn = k * n;
  // Minv is (kn x kn) row-major; block (l, n') starts at l*n*kn + n'*n with row stride kn.
  // Stage it as a dense n x n block in the part of R reserved for scratch.
  BlkAcc acc;
  acc.zero();
  const int tid = threadIdx.x;
  for (int np2 = 0; np2 < k; np2++) {
    const z2 *Rb = R + ((size_t)np2 * np + m) * p.es();
    // block (l, np2) of Minv: rows l*n.., columns np2*n.., row stride kn
    blk_mac(acc, s, n, Minv + (size_t)l * n * kn + (size_t)np2 * n, 0, Rb, 0, z2{1.0, 0.0}, kn, n);
  }
  z2 *g = p.gtv(l + 1, m);
  double d2 = blk_dist2(acc, n, g, red);
  blk_store(acc, n, g);
  if (tid == 0) atomicAdd(err, sqrt(d2));
}

// src/dyson_boot.h:78-88 and :181-192
__global__ void __launch_bounds__(256) kern_boot_retles(Pt p, int k, double *err) {
  __shared__ double red[256];
  const int blk = blockIdx.x, n = p.n;
  // enumerate (nn, l) pairs for ret: nn in 0..k, l in 0..k-nn; then les (l, nn): l in 0..k, nn in 0..l
  int cnt = 0;
  int kind = -1, A = 0, Bi = 0;
  for (int nn = 0; nn <= k && kind < 0; nn++)
    for (int l = 0; l <= k - nn; l++) {
      if (cnt == blk) { kind = 0; A = nn; Bi = l; break; }
      cnt++;
    }
  for (int l = 0; l <= k && kind < 0; l++)
    for (int nn = 0; nn <= l; nn++) {
      if (cnt == blk) { kind = 1; A = l; Bi = nn; break; }
      cnt++;
    }
  if (kind < 0) return;
  BlkAcc acc;
  acc.zero();
  z2 *g;
  if (kind == 0) {
    const int nn = A, l = Bi;
    if (nn == 0) {
      blk_diag(acc, n, z2{0.0, -1.0});
    } else {
      blk_axpy(acc, n, p.gtv(nn, p.np - 1), 0, z2{(double)p.sig, 0.0});
      blk_axpy(acc, n, p.gtv(nn, 0), 0, z2{-1.0, 0.0});
    }
    g = p.gr(nn + l, l);
  } else {
    const int l = A, nn = Bi;
    blk_axpy(acc, n, p.gtv(l - nn, 0), 1, z2{-1.0, 0.0});
    g = p.gl(nn, l);
  }
  double d2 = blk_dist2(acc, n, g, red);
  blk_store(acc, n, g);
  if (threadIdx.x == 0) atomicAdd(err, sqrt(d2));
}

__global__ void kern_mix_bs(const double *lb, const z2 *GM, int np, int nq, size_t es, int i, z2 *Bs) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)nq * es) return;
  int q = (int)(idx / es);
  size_t e = idx % es;
  const double *row = lb + ((size_t)i * nq + q) * np;
  z2 s = {0.0, 0.0};
  for (int k = 0; k < np; k++) s = zadd(s, zscale(GM[(size_t)k * es + e], row[k]));
  Bs[idx] = s;
}

__global__ void kern_mix_k(const double *la, const z2 *Bs, int np, int nq, size_t es, int i, z2 *K) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)np * es) return;
  int j = (int)(idx / es);
  size_t e = idx % es;
  z2 s = {0.0, 0.0};
  for (int q = 0; q < nq; q++) s = zadd(s, zscale(Bs[(size_t)q * es + e], la[((size_t)i * nq + q) * np + j]));
  K[((size_t)i * np + j) * es + e] = s;
}

// (for a context with FP32 Sigma storage the Sigma pointers address float2
// elements; SE_LAUNCH passes the kernels that read them the FP32 view)
Pt make_pt(Ctx &c, int T, double mu, double beta, double dt) {
  Pt p;
  p.Gret = c.G.ret.as<z2>(); p.Gles = c.G.les.as<z2>(); p.Gtv = c.G.tv.as<z2>(); p.Gmat = c.G.mat.as<z2>();
  p.Sret = c.S.ret.as<z2>(); p.Sles = c.S.les.as<z2>(); p.Stv = c.S.tv.as<z2>();
  p.hmf = c.hmf.as<z2>();
  p.n = c.nao; p.np = c.np; p.T = T; p.sig = c.sig;
  p.gring = c.gring; p.growlen = c.nt + 1;
  p.mu = mu; p.beta = beta; p.dt = dt;
  return p;
}

// launches kernel template K, which reads the Sigma triangles, for the
// context's Sigma storage: K<z2> with p, or K<float2> with its FP32 view
#define SE_LAUNCH(c, K, grid, block, shm, st, p, ...)                                         \
  do {                                                                                         \
    if ((c).s32)                                                                               \
      hipLaunchKernelGGL(K<float2>, grid, block, shm, st, as_s32(p), __VA_ARGS__);              \
    else                                                                                       \
      hipLaunchKernelGGL(K<z2>, grid, block, shm, st, p, __VA_ARGS__);                          \
  } while (0)
// the same for a group's launch over a device array of ChainArg (the two
// instances of ChainArgT have the same layout)
#define GRP_LAUNCH(s32, K, grid, block, shm, st, args, ...)                                                       \
  do {                                                                                                             \
    if (s32)                                                                                                       \
      hipLaunchKernelGGL(K<float2>, grid, block, shm, st,                                                          \
                         reinterpret_cast<const ChainArgT<float2> *>(args), __VA_ARGS__);                          \
    else                                                                                                           \
      hipLaunchKernelGGL(K<z2>, grid, block, shm, st, args, __VA_ARGS__);                                          \
  } while (0)

inline unsigned nblk(long long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

__global__ void kern_zero_words(unsigned *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// clears a small counter or error buffer with a kernel rather than a memset,
// so that the sequence stays valid inside a captured HIP graph
void zero_async(void *p, size_t bytes, hipStream_t s) {
  const int n = (int)((bytes + 3) / 4);
  hipLaunchKernelGGL(kern_zero_words, dim3(nblk(n)), dim3(256), 0, s, (unsigned *)p, n);
}

// (on stream st, c.stream if null, counting zero pivots in *info, the
// context's counter if null)
void gesv(Ctx &c, z2 *A, long long sA, z2 *B, long long sB, int N, int R, int nsys, hipStream_t st = nullptr,
          int *info = nullptr) {
  require(N <= 320, "linear system larger than 320");
  if (!st) st = c.stream;
  if (!info) info = c.dy_info.as<int>();
  zero_async(info, sizeof(int), st);
  if (N <= LDS_N && R <= LDS_N) {
    hipLaunchKernelGGL(kern_gesv_lds, dim3(nsys), dim3(256), 0, st, A, sA, B, sB, N, R, info);
  } else if (nsys == 1) {
    c.dy_pinv.ensure((size_t)N * sizeof(z2));
    z2 *pinv = c.dy_pinv.as<z2>();
    for (int col = 0; col < N; col++) {
      hipLaunchKernelGGL(kern_gj_pivot, dim3(1), dim3(256), 0, st, A, B, N, R, col, pinv, info);
      const long long work = (long long)(N - 1) * (N - col - 1 + R);
      hipLaunchKernelGGL(kern_gj_elim, dim3(nblk(work)), dim3(256), 0, st, A, B, N, R, col, pinv);
    }
    hipLaunchKernelGGL(kern_gj_scale, dim3(nblk((long long)N * R)), dim3(256), 0, st, B, N, R, pinv);
  } else {
    hipLaunchKernelGGL(kern_gesv, dim3(nsys), dim3(256), 0, st, A, sA, B, sB, N, R, info);
  }
  check_launch("gesv");
}

void update_minv(Ctx &c, int T, double mu, double dt, const WTab &w) {
  const size_t es = c.es;
  if (c.minv_mu != mu || c.minv_dt != dt) {
    invalidate_minv(c);
    c.minv_mu = mu; c.minv_dt = dt;
  }
  std::vector<int> js;
  for (int j = 0; j < T; j++) if (!c.minv_valid[j]) js.push_back(j);
  if (js.empty()) return;
  const int cnt = (int)js.size();
  c.dy_js.ensure(cnt * sizeof(int));
  NED_HIP(hipMemcpyAsync(c.dy_js.p, js.data(), cnt * sizeof(int), hipMemcpyHostToDevice, c.stream));
  c.dy_gA.ensure((size_t)cnt * es * sizeof(z2));
  c.dy_gB.ensure((size_t)cnt * es * sizeof(z2));
  Pt p = make_pt(c, T, mu, 0.0, dt);
  SE_LAUNCH(c, kern_build_mj, dim3(nblk((long long)cnt * es)), dim3(256), 0, c.stream, p, c.dy_js.as<int>(), cnt,
                     w.bd[0], w.Omega[w.k * (w.k + 1)], c.dy_gA.as<z2>(), c.dy_gB.as<z2>());
  check_launch("build mj");
  gesv(c, c.dy_gA.as<z2>(), es, c.dy_gB.as<z2>(), es, c.nao, c.nao, cnt);
  hipLaunchKernelGGL(kern_scatter_blocks, dim3(nblk((long long)cnt * es)), dim3(256), 0, c.stream, c.dy_gB.as<z2>(),
                     c.dy_js.as<int>(), cnt, es, c.Minv.as<z2>());
  check_launch("scatter minv");
  NED_HIP(hipStreamSynchronize(c.stream));  // js host vector goes out of scope
  for (int j : js) c.minv_valid[j] = 1;
}

// target number of (row, chunk) workgroups of the chunked history sums
constexpr int kPartBlocks = 12288;

int chunks_for(int rows, int terms) {
  // about 48 workgroups per CU, so that the last partial round of workgroups
  // is a small part of the launch, without chunks of fewer than 8 terms (at
  // nao = 64, T = 1000: 13 chunks a row, 0.4 ms a Dyson step faster than 9 or
  // 17)
  int want = (kPartBlocks + rows - 1) / std::max(rows, 1);
  want = std::min(want, std::max(1, terms / 8));
  return std::max(1, std::min(want, 64));
}

}  // namespace

void touch_time(Ctx &c, int t) {
  if (t >= 0 && t < (int)c.minv_valid.size()) c.minv_valid[t] = 0;
  if (t <= c.k) c.lesinv_valid = false;
  if (t < c.retinv_T) c.retinv_T = -1;
  // the transfer blocks of step T read times before T only
  for (int r = 0; r < 2; r++)
    if (t < c.tb_T[r]) c.tb_T[r] = -1;
}

void invalidate_minv(Ctx &c) {
  std::fill(c.minv_valid.begin(), c.minv_valid.end(), 0);
  c.retinv_T = -1;
  c.lesinv_valid = false;
  c.tb_T[0] = c.tb_T[1] = -1;
}

void set_matsubara(Ctx &c, double beta) {
  GLL g(c.ntau);
  const int np = c.np, nq = 2 * np;
  const size_t es = c.es;
  std::vector<double> la, lb;
  g.mixing_rows(beta, c.sig, la, lb);
  c.dy_la.ensure(la.size() * sizeof(double));
  c.dy_lb.ensure(lb.size() * sizeof(double));
  NED_HIP(hipMemcpyAsync(c.dy_la.p, la.data(), la.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
  NED_HIP(hipMemcpyAsync(c.dy_lb.p, lb.data(), lb.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
  c.mixK.ensure((size_t)np * np * es * sizeof(z2));
  c.dy_mixtmp.ensure((size_t)nq * es * sizeof(z2));
  for (int i = 0; i < np; i++) {
    hipLaunchKernelGGL(kern_mix_bs, dim3(nblk((long long)nq * es)), dim3(256), 0, c.stream, c.dy_lb.as<double>(),
                       c.G.mat.as<z2>(), np, nq, es, i, c.dy_mixtmp.as<z2>());
    hipLaunchKernelGGL(kern_mix_k, dim3(nblk((long long)np * es)), dim3(256), 0, c.stream, c.dy_la.as<double>(),
                       c.dy_mixtmp.as<z2>(), np, nq, es, i, c.mixK.as<z2>());
  }
  check_launch("mixing kernel");
  c.mass.ensure((size_t)np * np * sizeof(double));
  NED_HIP(hipMemcpyAsync(c.mass.p, g.mass.data(), g.mass.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
  NED_HIP(hipStreamSynchronize(c.stream));
  c.mix_beta = beta;
  c.have_mix = true;
}

namespace {

// The blocked right-looking recurrence with look-ahead, shared by the
// retarded and lesser phases, for one system or a group.  Panel p holds the
// links nn in [q0, q1).  Each link finishes block nn+1 and adds block nn to
// the later blocks of its own panel and of the next one (chain(q0, q1, q1x):
// the panel's links on stream s, the dependency chain, as one launch each or
// one launch for the panel; accumulator QA).  The bulk
// update -- panel p applied to every block beyond the next panel,
// panel(q0, q1, q1x, band) -- runs on the side stream sd into a second
// accumulator Q while the chain goes on, and the chain of panel p+2 waits for
// it (events evs).  A finishing link adds both sums.  Ends with s waiting for
// sd.
template <class Chain, class Panel>
void recurrence(int k, int T, int RP, int band_max, hipStream_t s, hipStream_t sd, hipEvent_t evm, hipEvent_t *evs,
                Chain chain, Panel panel) {
  int pi = 0;
  for (int q0 = k + 1; q0 < T; q0 += RP, pi++) {
    const int q1 = std::min(q0 + RP, T), q1x = std::min(q1 + RP, T);
    if (pi >= 2) NED_HIP(hipStreamWaitEvent(s, evs[pi & 1], 0));
    chain(q0, q1, q1x);
    if (q1x < T) {
      NED_HIP(hipEventRecord(evm, s));
      NED_HIP(hipStreamWaitEvent(sd, evm, 0));
      panel(q0, q1, q1x, T - q1x <= band_max);
      NED_HIP(hipEventRecord(evs[pi & 1], sd));
    }
  }
  NED_HIP(hipEventRecord(evm, sd));
  NED_HIP(hipStreamWaitEvent(s, evm, 0));
}

// blocks per panel; at least k, so that the chain links of a panel reach
// every block whose finishing terms they fold in (up to k+1 blocks ahead)
inline int panel_blocks(const Ctx &c) { return std::max(c.rpanel, c.k); }
inline int bands(const Ctx &c) { return (c.nao + 15) / 16; }  // 16-row / 16-column bands of a block

// One launch per panel for the dependency chains (kern_ret_chain /
// kern_les_chain) unless the context asks for one launch per link: the
// flags (one per block and band, for each recurrence), the ticket counters
// and the epoch are allocated and zeroed with the step's other scratch
// (chain_alloc); the epoch is bumped on the device for each recurrence
void chain_alloc(Ctx &c) {
  if (c.chain_mode != 2) return;
  const size_t nflag = 2 * (size_t)(c.nt + 1) * bands(c);
  if (c.dy_flags.bytes >= nflag * sizeof(unsigned)) return;
  c.dy_flags.ensure(nflag * sizeof(unsigned));
  c.dy_ctr.ensure(4 * sizeof(unsigned));
  hipLaunchKernelGGL(kern_zero_words, dim3(nblk((long long)nflag)), dim3(256), 0, c.stream, c.dy_flags.as<unsigned>(),
                     (int)nflag);
  hipLaunchKernelGGL(kern_zero_words, dim3(1), dim3(256), 0, c.stream, c.dy_ctr.as<unsigned>(), 4);
  check_launch("dyson chain flags");
}

bool chain_begin(Ctx &c) {
  if (c.chain_mode != 2) return false;
  hipLaunchKernelGGL(kern_chain_epoch, dim3(1), dim3(1), 0, c.stream, c.dy_ctr.as<unsigned>() + 2);
  check_launch("dyson chain epoch");
  return true;
}

// ---- phase 1: cached inverses, scratch, the mixing row on the third stream
// (src/dyson_step.h:130-169: it reads only Sigma and the G^rceil and G^M
// history, not the retarded row, so it runs beside the retarded recurrence
// with its own scratch), and the start-up of the retarded row
// (src/dyson_step.h:28-65) with the history sums' initial values
void dyson_begin(Ctx &c, int T, double mu, double beta, double dt) {
  require(c.have_mix && c.mix_beta == beta, "dyson_step: call ned_set_matsubara(beta) first");
  require(T > c.k && T <= c.nt, "dyson_step: tstp must be in (k, nt]");
  g_need(c, T - c.k - 1, T - 1, "dyson_step");
  g_wrote(c, T);
  const WTab w = make_wtab(c.k);
  const WTab *wd = dev_wtab(c);
  const int n = c.nao, k = c.k, kn = k * n, np = c.np;
  const size_t es = c.es;
  Pt p = make_pt(c, T, mu, beta, dt);
  hipStream_t s = c.stream;
  c.mark(4);
  update_minv(c, T, mu, dt, w);
  c.dy_err.ensure(4 * sizeof(double));
  double *err = c.dy_err.as<double>();
  zero_async(err, 4 * sizeof(double), s);
  // scratch sized for the last step (nt), so that it is allocated once and
  // no allocation (an implicit device synchronisation) falls between the
  // launches of a step
  c.dy_Q.ensure((size_t)(c.nt + 1) * es * sizeof(z2));
  c.dy_QA.ensure((size_t)(c.nt + 1) * es * sizeof(z2));
  c.dy_X.ensure((size_t)(c.nt + 1) * es * sizeof(z2));
  // the lesser history sums' partial sums: nch * (T + 1) <= kPartBlocks + T + 1 blocks
  c.dy_part.ensure((size_t)(kPartBlocks + c.nt + 1) * es * sizeof(z2));
  c.dy_Y.ensure((size_t)np * es * sizeof(z2));
  c.dy_M.ensure((size_t)kn * kn * sizeof(z2));
  c.dy_Qs.ensure((size_t)kn * n * sizeof(z2));
  c.dy_d2.ensure((size_t)(c.nt + 1) * bands(c) * sizeof(double));
  c.dy_retinv.ensure((size_t)kn * kn * sizeof(z2));
  c.dy_lesinv.ensure((size_t)kn * kn * sizeof(z2));
  c.dy_Xs.ensure((size_t)kn * n * sizeof(z2));
  chain_alloc(c);
  z2 *Q = c.dy_Q.as<z2>(), *M = c.dy_M.as<z2>(), *Qs = c.dy_Qs.as<z2>(), *QA = c.dy_QA.as<z2>();
  {
    const int nch = chunks_for(np, T + np);
    c.dy_tvpart.ensure((size_t)std::max(nch * np, 1) * es * sizeof(z2));
    c.dy_tvA.ensure(es * sizeof(z2));
    c.dy_tvB.ensure(es * sizeof(z2));
    c.dy_tvinfo.ensure(sizeof(int));
    c.dy_scr.ensure((size_t)std::max(np, c.nt + 1) * es * sizeof(z2));
    NED_HIP(hipEventRecord(c.ev_fork, s));
    NED_HIP(hipStreamWaitEvent(c.side2, c.ev_fork, 0));
    hipStream_t s2 = c.side2;
    SE_LAUNCH(c, kern_tv_matrix, dim3(1), dim3(256), 0, s2, p, wd, c.dy_tvA.as<z2>(), c.dy_tvB.as<z2>());
    gesv(c, c.dy_tvA.as<z2>(), 0, c.dy_tvB.as<z2>(), 0, n, n, 1, s2, c.dy_tvinfo.as<int>());
    SE_LAUNCH(c, kern_tv_rhs, dim3(np, nch), dim3(256), 0, s2, p, wd, c.mixK.as<z2>(), nch,
                       c.dy_tvpart.as<z2>());
    hipLaunchKernelGGL(kern_tv_fin, dim3(np), dim3(256), 0, s2, p, c.dy_tvpart.as<z2>(), nch, c.dy_tvB.as<z2>(),
                       c.dy_scr.as<z2>(), err + 1);
    check_launch("tv");
    NED_HIP(hipEventRecord(c.ev_tv, s2));
  }
  // ---- retarded start-up (src/dyson_step.h:7-127)
  z2 *Xs = c.dy_Xs.as<z2>();
  if (c.retinv_T != T) {
    SE_LAUNCH(c, kern_ret_start, dim3(nblk((long long)kn * kn)), dim3(256), 0, s, p, wd, 1, M,
                       c.dy_retinv.as<z2>());
    gesv(c, M, 0, c.dy_retinv.as<z2>(), 0, kn, kn, 1);
    c.retinv_T = T;
  }
  SE_LAUNCH(c, kern_ret_start, dim3(nblk((long long)kn * n)), dim3(256), 0, s, p, wd, 0, M, Qs);
  hipLaunchKernelGGL(kern_zmm, dim3(nblk((long long)kn * n)), dim3(256), 0, s, c.dy_retinv.as<z2>(), Qs, kn, n, Xs);
  check_launch("ret start");
  hipLaunchKernelGGL(kern_ret_start_store, dim3(k), dim3(256), 0, s, p, Xs, err + 0);
  SE_LAUNCH(c, kern_ret_qinit, dim3(T - k), dim3(256), 0, s, p, wd, Q, QA, c.Minv.as<z2>(), err + 0);
}

// The transfer blocks of the one-product links (kern_ret_tblk /
// kern_les_tblk; r = 0 retarded, 1 lesser): formed on the first call for a
// time step, kept for its corrector passes (their inputs, Sigma^R rows and
// inverses of the times before T, do not change between them; touch_time and
// invalidate_minv drop them when they do)
z2 *transfer_blocks(Ctx &c, int r, int T, double mu, double dt, const Pt &p, const WTab *wd) {
  const int RP = panel_blocks(c), nlink = T - c.k - 1;
  DevBuf &tb = r == 0 ? c.dy_tbr : c.dy_tbl;
  if (nlink < 1) return tb.as<z2>();
  if (c.tb_T[r] != T || c.tb_rp[r] != RP || c.tb_mu[r] != mu || c.tb_dt[r] != dt) {
    tb.ensure((size_t)(c.nt + 1) * RP * c.es * sizeof(z2));
    if (r == 0)
      SE_LAUNCH(c, kern_ret_tblk, dim3(RP, nlink), dim3(256), 0, c.stream, p, wd, RP, c.Minv.as<z2>(),
                tb.as<z2>());
    else
      SE_LAUNCH(c, kern_les_tblk, dim3(RP, nlink), dim3(256), 0, c.stream, p, wd, RP, c.Minv.as<z2>(),
                tb.as<z2>());
    check_launch("dyson transfer blocks");
    c.tb_T[r] = T;
    c.tb_rp[r] = RP;
    c.tb_mu[r] = mu;
    c.tb_dt[r] = dt;
  }
  return tb.as<z2>();
}

// ---- phase 2 (one system): the retarded recurrence
void dyson_ret_one(Ctx &c, int T, double mu, double beta, double dt) {
  const WTab *wd = dev_wtab(c);
  Pt p = make_pt(c, T, mu, beta, dt);
  const int nb = bands(c), RP = panel_blocks(c);
  z2 *Q = c.dy_Q.as<z2>(), *QA = c.dy_QA.as<z2>();
  const z2 *Minv = c.Minv.as<z2>();
  double *d2 = c.dy_d2.as<double>();
  const bool persist = chain_begin(c);
  const z2 *tb = c.chain_mode == 1 ? transfer_blocks(c, 0, T, mu, dt, p, wd) : nullptr;
  recurrence(c.k, T, RP, c.panel_band_max, c.stream, c.side, c.ev_main, c.ev_side,
             [&](int q0, int q1, int q1x) {
               if (persist) {
                 SE_LAUNCH(c, kern_ret_chain, dim3(nb * (q1x - q0)), dim3(256), 0, c.stream, p, wd, q0, q1, nb, QA,
                           Q, Minv, d2, c.dy_flags.as<unsigned>(), c.dy_ctr.as<unsigned>(),
                           c.dy_ctr.as<unsigned>() + 2, c.dy_err.as<double>() + 3);
                 return;
               }
               if (tb) {
                 hipLaunchKernelGGL(kern_ret_xform, dim3(nb, q1 - q0), dim3(256), 0, c.stream, p, q0, QA, Q, Minv);
                 for (int nn = q0; nn < q1; nn++)
                   SE_LAUNCH(c, kern_ret_link, dim3(nb, q1x - nn), dim3(256), 0, c.stream, p, wd, nn, q1, RP, QA, tb,
                             d2);
                 return;
               }
               for (int nn = q0; nn < q1; nn++)
                 SE_LAUNCH(c, kern_ret_step_band, dim3(nb, q1x - nn), dim3(256), 0, c.stream, p, wd, nn, QA, Q,
                           Minv, d2);
             },
             [&](int q0, int q1, int q1x, bool band) {
               if (band)
                 SE_LAUNCH(c, kern_ret_panel_band, dim3(T - q1x, nb), dim3(256), 0, c.side, p, wd, q0, q1,
                                    q1x + 1, Q);
               else
                 SE_LAUNCH(c, kern_ret_panel, dim3(T - q1x), dim3(256), 0, c.side, p, wd, q0, q1, q1x + 1, Q);
             });
  check_launch("ret history");
}

// ---- phase 3: the retarded update norm, the mixing row joined back, and
// the lesser column's history sums and start-up (src/dyson_step.h:171-251)
void dyson_mid(Ctx &c, int T, double mu, double beta, double dt) {
  const WTab *wd = dev_wtab(c);
  const int n = c.nao, k = c.k, kn = k * n, np = c.np, nb = bands(c);
  const size_t es = c.es;
  Pt p = make_pt(c, T, mu, beta, dt);
  hipStream_t s = c.stream;
  double *err = c.dy_err.as<double>();
  z2 *Q = c.dy_Q.as<z2>(), *X = c.dy_X.as<z2>(), *M = c.dy_M.as<z2>(), *Qs = c.dy_Qs.as<z2>();
  if (T > k + 1) hipLaunchKernelGGL(kern_band_err, dim3(1), dim3(256), 0, s, c.dy_d2.as<double>(), k + 2, T, nb,
                                    err + 0);
  NED_HIP(hipStreamWaitEvent(s, c.ev_tv, 0));
  hipLaunchKernelGGL(kern_les_Y, dim3(np, nblk((long long)es)), dim3(256), 0, s, p, c.mass.as<double>(),
                     c.dy_Y.as<z2>());
  int nch = chunks_for(T + 1, np + T + 1);
  require((long long)nch * (T + 1) <= (long long)kPartBlocks + c.nt + 1, "dyson_step: partial-sum buffer too small");
  SE_LAUNCH(c, kern_les_rhs, dim3(T + 1, nch), dim3(256), 0, s, p, wd, c.dy_Y.as<z2>(), nch, c.dy_part.as<z2>());
  long long nel = (long long)(T + 1) * es;
  hipLaunchKernelGGL(kern_sum_parts, dim3(nblk(nel)), dim3(256), 0, s, c.dy_part.as<z2>(), nch, nel, Q);
  const int build = c.lesinv_valid ? 0 : 1;
  hipLaunchKernelGGL(kern_les_x0, dim3(1), dim3(256), 0, s, p, X, err + 2);
  SE_LAUNCH(c, kern_les_start, dim3(nblk((long long)kn * (build ? kn : n))), dim3(256), 0, s, p, wd, build, Q,
                     X, M, c.dy_lesinv.as<z2>(), Qs);
  if (build) {
    gesv(c, M, 0, c.dy_lesinv.as<z2>(), 0, kn, kn, 1);
    c.lesinv_valid = true;
  }
  hipLaunchKernelGGL(kern_zmm, dim3(nblk((long long)kn * n)), dim3(256), 0, s, c.dy_lesinv.as<z2>(), Qs, kn, n,
                     X + es);
  SE_LAUNCH(c, kern_les_qinit, dim3(T - k), dim3(256), 0, s, p, wd, Q, c.dy_QA.as<z2>(), X, c.Minv.as<z2>());
  check_launch("les start");
}

// ---- phase 4 (one system): the lesser recurrence, the same schedule
void dyson_les_one(Ctx &c, int T, double mu, double beta, double dt) {
  const WTab *wd = dev_wtab(c);
  Pt p = make_pt(c, T, mu, beta, dt);
  const int nb = bands(c), RP = panel_blocks(c);
  z2 *Q = c.dy_Q.as<z2>(), *QA = c.dy_QA.as<z2>(), *X = c.dy_X.as<z2>();
  const z2 *Minv = c.Minv.as<z2>();
  const bool persist = chain_begin(c);
  const z2 *tb = c.chain_mode == 1 ? transfer_blocks(c, 1, T, mu, dt, p, wd) : nullptr;
  recurrence(c.k, T, RP, c.panel_band_max, c.stream, c.side, c.ev_main, c.ev_side,
             [&](int q0, int q1, int q1x) {
               if (persist) {
                 SE_LAUNCH(c, kern_les_chain, dim3(nb * (q1x - q0)), dim3(256), 0, c.stream, p, wd, q0, q1, nb, QA,
                           Q, X, Minv, c.dy_flags.as<unsigned>() + (size_t)(c.nt + 1) * nb, c.dy_ctr.as<unsigned>(),
                           c.dy_ctr.as<unsigned>() + 2, c.dy_err.as<double>() + 3);
                 return;
               }
               if (tb) {
                 const int qf = std::min(q1, T - 1);  // block T is left to the diagonal step
                 if (qf > q0)
                   hipLaunchKernelGGL(kern_les_xform, dim3(nb, qf - q0), dim3(256), 0, c.stream, p, q0, QA, Q, Minv);
                 for (int l = q0; l < q1; l++)
                   SE_LAUNCH(c, kern_les_link, dim3(nb, q1x - l), dim3(256), 0, c.stream, p, wd, l, q1, RP, QA, X,
                             tb);
                 return;
               }
               for (int l = q0; l < q1; l++)
                 SE_LAUNCH(c, kern_les_step_band, dim3(nb, q1x - l), dim3(256), 0, c.stream, p, wd, l, QA, Q, X,
                           Minv);
             },
             [&](int q0, int q1, int q1x, bool band) {
               if (band)
                 SE_LAUNCH(c, kern_les_panel_band, dim3(T - q1x, nb), dim3(256), 0, c.side, p, wd, q0, q1,
                                    q1x + 1, Q, X);
               else
                 SE_LAUNCH(c, kern_les_panel, dim3(T - q1x), dim3(256), 0, c.side, p, wd, q0, q1, q1x + 1, Q,
                                    X);
             });
  check_launch("les history");
}

// ---- phase 5: the Hermitian diagonal step (src/dyson_step.h:283-310) and the
// lesser column's store with its update norm
void dyson_end(Ctx &c, int T, double mu, double beta, double dt, double *err3) {
  const WTab *wd = dev_wtab(c);
  Pt p = make_pt(c, T, mu, beta, dt);
  hipStream_t s = c.stream;
  SE_LAUNCH(c, kern_les_diag, dim3(1), dim3(256), 0, s, p, wd, c.dy_Q.as<z2>(), c.dy_QA.as<z2>(),
                     c.dy_X.as<z2>(), c.dy_scr.as<z2>());
  hipLaunchKernelGGL(kern_les_store, dim3(T + 1), dim3(256), 0, s, p, c.dy_X.as<z2>(), c.dy_err.as<double>() + 2);
  check_launch("les");
  c.mark(5);
  if (err3) dyson_read_err(c, err3);
}

}  // namespace

void dyson_step(Ctx &c, int T, double mu, double beta, double dt, double *err3) {
  dyson_begin(c, T, mu, beta, dt);
  dyson_ret_one(c, T, mu, beta, dt);
  dyson_mid(c, T, mu, beta, dt);
  dyson_les_one(c, T, mu, beta, dt);
  dyson_end(c, T, mu, beta, dt, err3);
}

DysonGroup::DysonGroup(int device_) : device(device_) {
  NED_HIP(hipSetDevice(device));
  NED_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  NED_HIP(hipStreamCreateWithFlags(&sd, hipStreamNonBlocking));
  NED_HIP(hipEventCreateWithFlags(&evm, hipEventDisableTiming));
  for (hipEvent_t &e : evs) NED_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

DysonGroup::~DysonGroup() {
  (void)hipSetDevice(device);
  (void)hipStreamSynchronize(s);
  (void)hipStreamSynchronize(sd);
  for (hipEvent_t e : evr) (void)hipEventDestroy(e);
  for (hipEvent_t e : evs) (void)hipEventDestroy(e);
  if (evm) (void)hipEventDestroy(evm);
  if (s) (void)hipStreamDestroy(s);
  if (sd) (void)hipStreamDestroy(sd);
  if (host_args) (void)hipHostFree(host_args);
}

// The group's stream waits for every member's stream (and then the members
// wait for the group's) through one event per member.
static void group_join_in(DysonGroup &g, const std::vector<Ctx *> &cs) {
  for (size_t i = 0; i < cs.size(); i++) {
    NED_HIP(hipEventRecord(g.evr[i], cs[i]->stream));
    NED_HIP(hipStreamWaitEvent(g.s, g.evr[i], 0));
  }
}
static void group_join_out(DysonGroup &g, const std::vector<Ctx *> &cs) {
  NED_HIP(hipEventRecord(g.evm, g.s));
  for (Ctx *c : cs) NED_HIP(hipStreamWaitEvent(c->stream, g.evm, 0));
}

void dyson_step_group(DysonGroup &g, const std::vector<Ctx *> &cs, int T, const double *mu, const double *beta,
                      const double *dt, double *err3) {
  const int S = (int)cs.size();
  require(S > 0, "dyson group: no systems");
  const Ctx &c0 = *cs[0];
  int nbmax = 0;
  for (Ctx *c : cs) {
    require(c->device == g.device, "dyson group: every system must be on the group's device");
    require(c->k == c0.k && c->rpanel == c0.rpanel && c->panel_band_max == c0.panel_band_max,
            "dyson group: the systems must share k and the recurrence schedule");
    require(c->s32 == c0.s32, "dyson group: the systems must share the Sigma storage precision");
    nbmax = std::max(nbmax, bands(*c));
  }
  const bool s32 = c0.s32;
  while ((int)g.evr.size() < S) {
    hipEvent_t e;
    NED_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    g.evr.push_back(e);
  }
  for (int i = 0; i < S; i++) dyson_begin(*cs[i], T, mu[i], beta[i], dt[i]);
  // the members' chain operands, copied from pinned host memory (the previous
  // copy has completed: every pass ends with a read of the update norms)
  const size_t bytes = (size_t)S * sizeof(ChainArg);
  if (g.host_cap < bytes) {
    if (g.host_args) NED_HIP(hipHostFree(g.host_args));
    g.host_args = nullptr;
    NED_HIP(hipHostMalloc(&g.host_args, bytes));
    g.host_cap = bytes;
  }
  g.args.ensure(bytes);
  ChainArg *ha = reinterpret_cast<ChainArg *>(g.host_args);
  for (int i = 0; i < S; i++) {
    Ctx &c = *cs[i];
    ha[i] = ChainArg{make_pt(c, T, mu[i], beta[i], dt[i]), c.dy_QA.as<z2>(), c.dy_Q.as<z2>(), c.dy_X.as<z2>(),
                     c.Minv.as<z2>(), c.dy_d2.as<double>(), bands(c)};
  }
  group_join_in(g, cs);
  NED_HIP(hipMemcpyAsync(g.args.p, g.host_args, bytes, hipMemcpyHostToDevice, g.s));
  const ChainArg *args = g.args.as<ChainArg>();
  const WTab *wd = dev_wtab(*cs[0]);
  recurrence(c0.k, T, panel_blocks(c0), c0.panel_band_max, g.s, g.sd, g.evm, g.evs,
             [&](int q0, int q1, int q1x) {
               for (int nn = q0; nn < q1; nn++)
                 GRP_LAUNCH(s32, kern_ret_step_band_grp, dim3(nbmax, q1x - nn, S), dim3(256), 0, g.s, args, wd, nn);
             },
             [&](int q0, int q1, int q1x, bool band) {
               if (band)
                 GRP_LAUNCH(s32, kern_ret_panel_band_grp, dim3(T - q1x, nbmax, S), dim3(256), 0, g.sd, args, wd,
                                    q0, q1, q1x + 1);
               else
                 GRP_LAUNCH(s32, kern_ret_panel_grp, dim3(T - q1x, S), dim3(256), 0, g.sd, args, wd, q0, q1,
                                    q1x + 1);
             });
  check_launch("group ret history");
  group_join_out(g, cs);
  for (int i = 0; i < S; i++) dyson_mid(*cs[i], T, mu[i], beta[i], dt[i]);
  group_join_in(g, cs);
  recurrence(c0.k, T, panel_blocks(c0), c0.panel_band_max, g.s, g.sd, g.evm, g.evs,
             [&](int q0, int q1, int q1x) {
               for (int l = q0; l < q1; l++)
                 GRP_LAUNCH(s32, kern_les_step_band_grp, dim3(nbmax, q1x - l, S), dim3(256), 0, g.s, args, wd, l);
             },
             [&](int q0, int q1, int q1x, bool band) {
               if (band)
                 GRP_LAUNCH(s32, kern_les_panel_band_grp, dim3(T - q1x, nbmax, S), dim3(256), 0, g.sd, args, wd,
                                    q0, q1, q1x + 1);
               else
                 GRP_LAUNCH(s32, kern_les_panel_grp, dim3(T - q1x, S), dim3(256), 0, g.sd, args, wd, q0, q1,
                                    q1x + 1);
             });
  check_launch("group les history");
  group_join_out(g, cs);
  for (int i = 0; i < S; i++) dyson_end(*cs[i], T, mu[i], beta[i], dt[i], err3 ? err3 + 3 * i : nullptr);
}

void dyson_read_err(Ctx &c, double *err3) {
  double h[4];
  NED_HIP(hipMemcpyAsync(h, c.dy_err.p, 4 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  NED_HIP(hipStreamSynchronize(c.stream));
  if (h[3] != 0.0) throw std::runtime_error("dyson_step: a wait of the one-launch-per-panel dependency chain passed its bound");
  err3[0] = h[0]; err3[1] = h[1]; err3[2] = h[2];
}

void extrapolate(Ctx &c, int T) {
  require(T > c.k && T <= c.nt, "extrapolate: tstp must be in (k, nt]");
  g_need(c, T - c.k - 1, T - 1, "extrapolate");
  g_wrote(c, T);
  const WTab w = make_wtab(c.k);
  const WTab *wd = dev_wtab(c);
  Pt p = make_pt(c, T, 0, 0, 0);
  const size_t es = c.es;
  hipLaunchKernelGGL(kern_extrap, dim3(nblk((long long)c.np * es)), dim3(256), 0, c.stream, p, wd, 0);
  hipLaunchKernelGGL(kern_extrap, dim3(nblk((long long)(T + 1) * es)), dim3(256), 0, c.stream, p, wd, 1);
  // the les column needs the new G^rceil(T,0) (row 0) -- launched after tv
  hipLaunchKernelGGL(kern_extrap, dim3(nblk((long long)(T + 1) * es)), dim3(256), 0, c.stream, p, wd, 2);
  check_launch("extrapolate");
}

void hf_decomp(Ctx &c, int T, const double *h0_host) {
  hf_upload_h0(c, h0_host);
  hf_kernels(c, T);
  NED_HIP(hipStreamSynchronize(c.stream));  // h0_host may be released by the caller
  touch_time(c, T);
}

// workgroup rows of kern_hf_h: about 1024 workgroups at nao = 64, at most
// one term each for tiny systems
int hf_splits(const Ctx &c) {
  const long long wg = (long long)(c.es + 255) / 256;
  return (int)std::max(1LL, std::min((long long)(c.naux + c.nao * c.nao), (1024 + wg - 1) / wg));
}

void hf_upload_h0(Ctx &c, const double *h0_host) {
  const size_t es = c.es;
  c.dy_h0.ensure(es * sizeof(double));
  c.dy_xq.ensure((size_t)c.naux * sizeof(z2));
  c.dy_hfpart.ensure((size_t)hf_splits(c) * es * sizeof(z2));
  NED_HIP(hipMemcpyAsync(c.dy_h0.p, h0_host, es * sizeof(double), hipMemcpyHostToDevice, c.stream));
}

// the mean field of step T from the h0 uploaded last (kernels only)
void hf_kernels(Ctx &c, int T) {
  require(c.have_coulomb, "hf: Coulomb tensor not set");
  require(T >= 0 && T <= c.nt, "hf: tstp out of range");
  g_need(c, T, T, "hf");
  const size_t es = c.es;
  Pt p = make_pt(c, T, 0, 0, 0);
  hipLaunchKernelGGL(kern_hf_x, dim3(c.naux), dim3(256), 0, c.stream, p, c.Bq.as<double>(), c.dy_xq.as<z2>());
  const int S = hf_splits(c);
  require(c.dy_hfpart.bytes >= (size_t)S * es * sizeof(z2), "hf: partial-sum buffer not allocated (upload h0 first)");
  hipLaunchKernelGGL(kern_hf_h, dim3(nblk((long long)es), S), dim3(256), 0, c.stream, p, c.Bq.as<double>(),
                     c.U64.as<double>(), c.dy_xq.as<z2>(), c.naux, c.dy_hfpart.as<z2>());
  hipLaunchKernelGGL(kern_hf_sum, dim3(nblk((long long)es)), dim3(256), 0, c.stream, p, c.dy_h0.as<double>(),
                     c.dy_hfpart.as<z2>(), S, c.hmf.as<z2>());
  check_launch("hf");
}

double dyson_start(Ctx &c, double mu, double beta, double dt) {
  require(c.have_mix && c.mix_beta == beta, "dyson_start: call ned_set_matsubara(beta) first");
  require(c.nt >= c.k, "dyson_start: nt must be >= k");
  g_need(c, 0, c.k, "dyson_start");
  const WTab w = make_wtab(c.k);
  const WTab *wd = dev_wtab(c);
  const int n = c.nao, k = c.k, kn = k * n, np = c.np;
  const size_t es = c.es;
  Pt p = make_pt(c, 0, mu, beta, dt);
  hipStream_t s = c.stream;
  c.dy_err.ensure(4 * sizeof(double));
  double *err = c.dy_err.as<double>();
  zero_async(err, 4 * sizeof(double), s);
  hipLaunchKernelGGL(kern_boot_tv0, dim3(np), dim3(256), 0, s, p, err);
  c.dy_M.ensure((size_t)kn * kn * sizeof(z2));
  c.dy_bootI.ensure((size_t)kn * kn * sizeof(z2));
  SE_LAUNCH(c, kern_boot_matrix, dim3(nblk((long long)kn * kn)), dim3(256), 0, s, p, wd, c.dy_M.as<z2>(),
                     c.dy_bootI.as<z2>());
  gesv(c, c.dy_M.as<z2>(), 0, c.dy_bootI.as<z2>(), 0, kn, kn, 1);
  c.dy_part.ensure((size_t)k * np * es * sizeof(z2));
  SE_LAUNCH(c, kern_boot_rhs, dim3(np, k), dim3(256), 0, s, p, wd, c.mixK.as<z2>(), c.dy_part.as<z2>());
  hipLaunchKernelGGL(kern_boot_apply, dim3(np, k), dim3(256), 0, s, p, k, c.dy_bootI.as<z2>(), c.dy_part.as<z2>(), err);
  int npairs = (k + 1) * (k + 2);  // ret pairs + les pairs
  hipLaunchKernelGGL(kern_boot_retles, dim3(npairs), dim3(256), 0, s, p, k, err);
  check_launch("dyson_start");
  double h;
  NED_HIP(hipMemcpyAsync(&h, err, sizeof(double), hipMemcpyDeviceToHost, s));
  NED_HIP(hipStreamSynchronize(s));
  invalidate_minv(c);
  return h;
}

void energy(Ctx &c, int T, const double *h0_host, double beta, double dt, double *out3) {
  require(c.have_mix && c.mix_beta == beta, "energy: call ned_set_matsubara(beta) first");
  const WTab w = make_wtab(c.k);
  const WTab *wd = dev_wtab(c);
  const int top = T >= c.k ? T : c.k, np = c.np;
  g_need(c, T, top, "energy");
  const int nb = top + 1 + np;
  c.dy_en.ensure((size_t)2 * nb * sizeof(double));
  Pt p = make_pt(c, T, 0, beta, dt);
  SE_LAUNCH(c, kern_energy, dim3(nb), dim3(256), 0, c.stream, p, wd, top, c.mass.as<double>(), c.dy_en.as<double>());
  check_launch("energy");
  std::vector<double> h(2 * nb);
  std::vector<z2> gl(c.es), hm(c.es);
  NED_HIP(hipMemcpyAsync(h.data(), c.dy_en.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  NED_HIP(hipMemcpyAsync(gl.data(), c.G.les.as<z2>() + (g_row(c, T) + T) * c.es, c.es * sizeof(z2),
                         hipMemcpyDeviceToHost, c.stream));
  NED_HIP(hipMemcpyAsync(hm.data(), c.hmf.as<z2>() + (size_t)T * c.es, c.es * sizeof(z2), hipMemcpyDeviceToHost,
                         c.stream));
  NED_HIP(hipStreamSynchronize(c.stream));
  double ry = 0;
  for (int i = 0; i < nb; i++) ry += h[2 * i + 1];
  // E = Re(i sig/2 (res1 + res2))
  const double e_corr = -0.5 * c.sig * ry;
  const int n = c.nao;
  double ekin = 0, ntr = 0;
  for (int a = 0; a < n; a++) {
    for (int b = 0; b < n; b++) {
      z2 g = gl[a * n + b];
      z2 rho = {-c.sig * g.y, c.sig * g.x};           // rho_ab = i sig G^<_ab
      z2 hb = hm[b * n + a];                            // (hmf + h0)^T_ab
      hb.x += h0_host[b * n + a];
      ekin += rho.x * hb.x - rho.y * hb.y;
    }
    z2 g = gl[a * n + a];
    ntr += -c.sig * g.y;
  }
  out3[0] = ekin;
  out3[1] = 2 * e_corr;
  out3[2] = 2 * ntr;
}

}  // namespace ned
