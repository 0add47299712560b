// Second-Born self-energy from the Cholesky-decomposed interaction, one time
// slice T at a time: the contract of molGF2SolverDecomp::solve
// (src/mol2bSolver_decomp_impl.h:412-429), restructured for the matrix cores.
//
// Every component is an instance of one functional of two nao x nao matrices,
//   F[A,B]_ij = sum_{lnpkqm} (2 U_{il,np} U_{kj,qm} - U_{il,np} U_{qj,km}) A_lk B_mn A_pq
// with U_{ij,kl} = sum_Q V^Q_ij V^Q_kl (the index pattern of solve_loop,
// src/mol2bSolver_decomp_impl.h:459-467, 585-593).  Writing W^a = V^a A,
//   bubble   2 sum_a W^a (sum_b Pi_ab V^b),  Pi_ab = tr(V^a A V^b B)
//   exchange  -sum_a W^a D^a,  D^a_kj = sum_{mq} U_{km,qj} (B W^a)_mq
// The exchange is reassociated through U, so its n^4 naux work is a single
// real GEMM [(item, plane, a) x (m,q)] x U'[(m,q) x (k,j)] that streams the
// whole slice through the MFMA units at once (the reference does two n^4 naux
// GEMMs per time point, lines 206-226).  At nao = 64 that GEMM is split by the
// pair symmetry of U' into two of half the flops (kern_build_Usym).
//
// Items of slice T (L = G^<(t,T), R = G^R(T,t)):
//   "<" t:  A = G^<(T,t) = -L^H,      B = G^>(t,T) = L - R^H   -> Sigma^<(T,t)
//   ">" t:  A = G^>(T,t) = R - L^H,   B = G^<(t,T) = L         -> Sigma^>(T,t)
//   tv tau: A = G^rceil(T,tau),       B = G^lceil(tau,T) = -sig G^rceil(T,beta-tau)^H
// and then Sigma^R(T,t) = Sigma^>(T,t) - Sigma^<(T,t) (Langreth; equal to the
// four-term expression of lines 620-624) and Sigma^<(t,T) = -Sigma^<(T,t)^H,
// which holds for symmetric V^Q (real orbitals; checked at upload).
#include "ctx.h"
#include "cgemm64.h"
#include <cmath>
#include <cstdlib>
#include <algorithm>
#include <map>
#include <mutex>
#include <string>

namespace ned {

size_t tri(int t) { return (size_t)t * (t + 1) / 2; }
__device__ __forceinline__ size_t dtri_s(int t) { return (size_t)t * (t + 1) / 2; }
void require(bool cond, const char *msg) {
  if (!cond) throw std::invalid_argument(msg);
}

namespace {

__global__ void kern_build_U(const double *Bq, int n, int a, double *U64, float *U32) {
  // U'[(m,q)][(k,j)] = sum_Q B[Q][k][m] B[Q][q][j]
  long long n2 = (long long)n * n;
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n2 * n2) return;
  int row = (int)(idx / n2), col = (int)(idx % n2);
  int m = row / n, q = row % n, k = col / n, j = col % n;
  double s = 0;
  for (int Q = 0; Q < a; Q++) s += Bq[(long long)Q * n2 + k * n + m] * Bq[(long long)Q * n2 + q * n + j];
  U64[idx] = s;
  U32[idx] = (float)s;
}

// Baug[(k,j)][c] = -U'[c][(k,j)] for c < n2, 2 B[c-n2][k][j] for c >= n2,
// split into bf16 hi / lo planes (row stride n2 + a).  With swap_rows the
// rows are ordered (j,k) instead, so that the GEMM writes each H^a transposed.
__global__ void kern_build_Uaug(const float *U, const float *Bq, int n, int a, int swap_rows, uint16_t *hi,
                                uint16_t *lo) {
  const int n2 = n * n;
  const long long ka = n2 + a;
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)n2 * ka) return;
  const long long rr = idx / ka, col = idx % ka;
  const long long r = swap_rows ? (rr % n) * n + rr / n : rr;
  const float v = col < n2 ? -U[col * n2 + r] : 2.0f * Bq[(col - n2) * n2 + r];
  const uint16_t h = f2bf(v);
  hi[idx] = h;
  lo[idx] = f2bf(v - bf2f(h));
}

// Pair table of the split exchange: entries [0, ns) are the pairs m <= q and
// [ns, ns + na) the pairs m < q, each row-major over the upper triangle,
// packed m << 8 | q.
__global__ void kern_sym_pairs(int n, uint16_t *pairs) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x, ns = n * (n + 1) / 2, na = n * (n - 1) / 2;
  if (t >= ns + na) return;
  const int strict = t >= ns ? 1 : 0;
  int r = strict ? t - ns : t, m = 0;
  while (r >= n - m - strict) { r -= n - m - strict; m++; }
  pairs[t] = (uint16_t)(m << 8 | (m + strict + r));
}

// Right operands of the split exchange.  U' is invariant under swapping both
// pairs, U'[(q,m)][(j,k)] = U'[(m,q)][(k,j)] (B symmetric), so with
//   Us[(m,q)][(k,j)] = -(U'[(m,q)][(k,j)] + U'[(m,q)][(j,k)]) / 2  (even in m,q)
//   Ua[(m,q)][(k,j)] = -(U'[(m,q)][(k,j)] - U'[(m,q)][(j,k)]) / 2  (odd in m,q)
// the exchange -C U' splits into Cs = xs Us over pairs m <= q (xs = X_mq + X_qm,
// X_mm on the diagonal) and Ca = xa Ua over pairs m < q (xa = X_mq - X_qm),
// with (-C U')_kj = Cs_kj + Ca_kj and (-C U')_jk = Cs_kj - Ca_kj for k < j.
// Rows are the output pairs (k,j): for both operands the na pairs k < j in the
// table's order, so that Cs and Ca of one pair share a column of the outputs,
// and for the symmetric one the diagonal (k,k) after them.  Columns are the K
// pairs; the symmetric operand carries 2 B[b][k][j] in columns ns + b (Pi V is
// symmetric in (k,j)).  Rows and columns past the pairs are zero.  Split into
// bf16 hi / lo.
__global__ void kern_build_Usym(const double *U, const double *Bq, const uint16_t *pairs, int n, int a, int anti,
                                int rows, int cols, uint16_t *hi, uint16_t *lo) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)rows * cols) return;
  const int r = (int)(idx / cols), c = (int)(idx % cols);
  const int n2 = n * n, ns = n * (n + 1) / 2, na = n * (n - 1) / 2, np = anti ? na : ns, off = anti ? ns : 0;
  double v = 0;
  if (r < np) {
    const int rk = r < na ? pairs[ns + r] : (r - na) * 257, k = rk >> 8, j = rk & 255;
    if (c < np) {
      const int mq = pairs[off + c], m = mq >> 8, q = mq & 255;
      const double *Ur = U + (size_t)(m * n + q) * n2;
      v = -0.5 * (Ur[k * n + j] + (anti ? -1.0 : 1.0) * Ur[j * n + k]);
    } else if (!anti && c < ns + a) {
      v = 2.0 * Bq[(size_t)(c - ns) * n2 + k * n + j];
    }
  }
  const float f = (float)v;
  const uint16_t h = f2bf(f);
  hi[idx] = h;
  lo[idx] = f2bf(f - bf2f(h));
}

// The split exchange of the FP64 / FP32 path (and of bf16x3 at shapes without
// the split-bf16 GEMMs).  Pair indices follow the table of kern_sym_pairs:
// pairs k <= j row-major over the upper triangle, then pairs k < j.
__device__ __forceinline__ int pair_sym(int n, int k, int j) { return k * n - k * (k - 1) / 2 + (j - k); }
__device__ __forceinline__ int pair_anti(int n, int k, int j) { return k * (n - 1) - k * (k - 1) / 2 + (j - k - 1); }

// Right operand, (ns + a + na) x ... as two row-major blocks: first the
// (ns + a) x ns block of the symmetric part, K rows (m,q) pairs m <= q then the
// Pi columns b, N columns the output pairs (k,j), k <= j:
//   -(U'[(m,q)][(k,j)] + U'[(m,q)][(j,k)]) / 2,   2 B[b][k][j];
// then the na x na block of the antisymmetric part over pairs m < q, k < j:
//   -(U'[(m,q)][(k,j)] - U'[(m,q)][(j,k)]) / 2.
template <typename T>
__global__ void kern_build_Usym_gen(const double *U, const double *Bq, const uint16_t *pairs, int n, int a, T *out) {
  const long long ns = n * (n + 1) / 2, na = n * (n - 1) / 2, n2 = (long long)n * n;
  const long long nsym = (ns + a) * ns;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nsym + na * na) return;
  double v;
  if (idx < nsym) {
    const long long c = idx / ns, o = idx % ns;
    const int kj = pairs[o], k = kj >> 8, j = kj & 255;
    if (c < ns) {
      const int mq = pairs[c], m = mq >> 8, q = mq & 255;
      const double *Ur = U + (size_t)(m * n + q) * n2;
      v = -0.5 * (Ur[k * n + j] + Ur[j * n + k]);
    } else {
      v = 2.0 * Bq[(size_t)(c - ns) * n2 + k * n + j];
    }
  } else {
    const long long c = (idx - nsym) / na, o = (idx - nsym) % na;
    const int kj = pairs[ns + o], k = kj >> 8, j = kj & 255;
    const int mq = pairs[ns + c], m = mq >> 8, q = mq & 255;
    const double *Ur = U + (size_t)(m * n + q) * n2;
    v = -0.5 * (Ur[k * n + j] - Ur[j * n + k]);
  }
  out[idx] = (T)v;
}

// rows r = (item, plane, x) of C^x (n x n each, X) and Pi (planar per item)
// -> XF rows [X_mq + X_qm (m <= q; X_mm) | Pi[plane][x][b] | X_mq - X_qm (m < q)]
template <typename T>
__global__ void kern_fold_pairs(const T *X, const T *Pi, const uint16_t *pairs, int n, int a, long long rows, T *XF) {
  const int ns = n * (n + 1) / 2, na = n * (n - 1) / 2;
  const long long kx = ns + a + na, n2 = (long long)n * n;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * kx) return;
  const long long r = idx / kx;
  const int c = (int)(idx % kx);
  const T *xr = X + r * n2;
  T v;
  if (c < ns) {
    const int mq = pairs[c], m = mq >> 8, q = mq & 255;
    v = m == q ? xr[m * n + m] : xr[m * n + q] + xr[q * n + m];
  } else if (c < ns + a) {
    const long long p = r / (2 * a), pl = (r / a) % 2, x = r % a;
    v = Pi[(p * 2 + pl) * a * a + x * a + (c - ns)];
  } else {
    const int mq = pairs[ns + (c - ns - a)], m = mq >> 8, q = mq & 255;
    v = xr[m * n + q] - xr[q * n + m];
  }
  XF[idx] = v;
}

// H rows (n x n, row-major (k,j)) from Y rows [Cs (ns) | Ca (na)]:
// H_kj = Cs + Ca, H_jk = Cs - Ca (k < j), H_kk = Cs
template <typename T>
__global__ void kern_unfold_pairs(const T *Y, int n, long long rows, T *H) {
  const int ns = n * (n + 1) / 2, na = n * (n - 1) / 2;
  const long long n2 = (long long)n * n;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * n2) return;
  const long long r = idx / n2;
  const int e = (int)(idx % n2), k = e / n, j = e % n;
  const T *y = Y + r * (ns + na);
  T v;
  if (k == j) v = y[pair_sym(n, k, k)];
  else if (k < j) v = y[pair_sym(n, k, j)] + y[ns + pair_anti(n, k, j)];
  else v = y[pair_sym(n, j, k)] - y[ns + pair_anti(n, j, k)];
  H[idx] = v;
}

// Pi of every item from the real product Pf = [Wr; Wi] [Zr; Zi]^T (2a x 2a):
// Pi_r = Pf_rr - Pf_ii, Pi_i = Pf_ri + Pf_ir, split into bf16 and written to
// rows (item, plane, a), columns col0 + b of the exchange GEMM's operand A
__global__ void kern_pi_combine(const float *Pf, int a, int P, uint16_t *hi, uint16_t *lo, long long ldh, int col0) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long a2 = (long long)a * a;
  if (idx >= (long long)P * 2 * a2) return;
  const int p = (int)(idx / (2 * a2));
  const long long rem = idx % (2 * a2);
  const int pl = (int)(rem / a2), x = (int)((rem % a2) / a), b = (int)(rem % a);
  const float *F = Pf + (size_t)p * 4 * a2;
  const long long w2 = 2LL * a;
  float v;
  if (pl == 0) v = F[x * w2 + b] - F[(a + x) * w2 + a + b];
  else v = F[x * w2 + a + b] + F[(a + x) * w2 + b];
  const size_t o = ((size_t)p * 2 * a + (size_t)pl * a + x) * ldh + col0 + b;
  const uint16_t h = f2bf(v);
  hi[o] = h;
  lo[o] = f2bf(v - bf2f(h));
}

// the same four columns b..b+3 at a time (a, ldh and col0 multiples of 4):
// float4 reads and 8-byte split stores
__global__ void kern_pi_combine4(const float *Pf, int a, int P, uint16_t *hi, uint16_t *lo, long long ldh, int col0) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long a2 = (long long)a * a, q2 = a2 / 4;
  if (idx >= (long long)P * 2 * q2) return;
  const int p = (int)(idx / (2 * q2));
  const long long rem = idx % (2 * q2);
  const int pl = (int)(rem / q2), x = (int)((rem % q2) / (a / 4)), b = 4 * (int)(rem % (a / 4));
  const float *F = Pf + (size_t)p * 4 * a2;
  const long long w2 = 2LL * a;
  const float4 u = *reinterpret_cast<const float4 *>(F + x * w2 + (pl ? a + b : b));
  const float4 t = *reinterpret_cast<const float4 *>(F + (a + x) * w2 + (pl ? b : a + b));
  const float v[4] = {pl ? u.x + t.x : u.x - t.x, pl ? u.y + t.y : u.y - t.y, pl ? u.z + t.z : u.z - t.z,
                      pl ? u.w + t.w : u.w - t.w};
  uint16_t h[4], l[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    h[q] = f2bf(v[q]);
    l[q] = f2bf(v[q] - bf2f(h[q]));
  }
  const size_t o = ((size_t)p * 2 * a + (size_t)pl * a + x) * ldh + col0 + b;
  *reinterpret_cast<uint2 *>(hi + o) =
      uint2{(uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16)};
  *reinterpret_cast<uint2 *>(lo + o) =
      uint2{(uint32_t)l[0] | ((uint32_t)l[1] << 16), (uint32_t)l[2] | ((uint32_t)l[3] << 16)};
}

void pi_combine(const float *Pf, int a, int P, uint16_t *hi, uint16_t *lo, long long ldh, int col0, hipStream_t s) {
  const long long tot = (long long)P * 2 * a * a;
  if (a % 4 == 0 && ldh % 4 == 0 && col0 % 4 == 0)
    hipLaunchKernelGGL(kern_pi_combine4, dim3((unsigned)((tot / 4 + 255) / 256)), dim3(256), 0, s, Pf, a, P, hi, lo,
                       ldh, col0);
  else
    hipLaunchKernelGGL(kern_pi_combine, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, Pf, a, P, hi, lo, ldh,
                       col0);
}

__global__ void kern_d2f(const double *x, float *y, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}

template <typename T>
__global__ void kern_items(const z2 *Gret, const z2 *Gles, const z2 *Gtv, int T_, size_t rowT, int n, int np, int sig,
                           int p0, int P, T *Ap, T *Bp, int transA, T *Btp) {
  const int n2 = n * n;
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)P * n2) return;
  int p = (int)(idx / n2), e = (int)(idx % n2), i = e / n, k = e % n;
  int g = p0 + p;
  z2 A, B;
  if (g <= 2 * T_ + 1) {  // row T_ of G^R and G^< starts at block rowT
    int t = g <= T_ ? g : g - T_ - 1;
    const z2 *L = Gles + (rowT + t) * n2, *R = Gret + (rowT + t) * n2;
    z2 Lik = L[i * n + k], Lki = L[k * n + i], Rik = R[i * n + k], Rki = R[k * n + i];
    if (g <= T_) {  // "<"
      A = {-Lki.x, Lki.y};                         // -conj(L_ki)
      B = {Lik.x - Rki.x, Lik.y + Rki.y};          // L_ik - conj(R_ki)
    } else {        // ">"
      A = {Rik.x - Lki.x, Rik.y + Lki.y};          // R_ik - conj(L_ki)
      B = Lik;
    }
  } else {
    int tau = g - 2 * T_ - 2;
    const z2 *Ga = Gtv + ((size_t)T_ * np + tau) * n2;
    const z2 *Gb = Gtv + ((size_t)T_ * np + (np - 1 - tau)) * n2;
    A = Ga[i * n + k];
    z2 bk = Gb[k * n + i];
    B = {-(double)sig * bk.x, (double)sig * bk.y};  // -sig conj
  }
  // A as stored, or transposed (A^T[k][i] = A[i][k]); B as stored, and
  // optionally also transposed
  T *a = Ap + (size_t)p * 2 * n2, *b = Bp + (size_t)p * 2 * n2;
  const int ea = transA ? k * n + i : e;
  a[ea] = (T)A.x; a[n2 + ea] = (T)A.y;
  b[e] = (T)B.x; b[n2 + e] = (T)B.y;
  if (Btp) {
    T *bt = Btp + (size_t)p * 2 * n2;
    bt[k * n + i] = (T)B.x; bt[n2 + k * n + i] = (T)B.y;
  }
}

// Sigma of the items p0 .. p0+P-1 of slice T_ into the contour storage: each
// item's value is the sum of its S partial products (part[(p S + s) 2 n2 +
// plane n2 + e]).  "<" item t writes Sigma^<(t,T) = -Sigma^<(T,t)^H and
// Sigma^R(T,t) = Sigma^>(T,t) - Sigma^<(T,t) when its ">" partner (item
// t + T_ + 1) is in the same chunk, -Sigma^<(T,t) otherwise, to which the
// partner's chunk, issued later on the same stream, then adds Sigma^>(T,t)
// (the same rounding: a - b is a + (-b)); tv items write Sigma^rceil(T,tau).
template <typename T, typename SE>
__global__ void kern_sum_scatter(const T *part, int S, int p0, int P, int T_, int n, int np, SE *Sret, SE *Sles,
                                 z2 *Stv) {
  const int n2 = n * n;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)P * n2) return;
  const int p = (int)(idx / n2), e = (int)(idx % n2), i = e / n, k = e % n;
  const int g = p0 + p, nles = T_ + 1;
  const size_t per = 2 * (size_t)n2, rowT = dtri_s(T_);
  auto item = [&](int q) {
    T re = 0, im = 0;
    for (int s = 0; s < S; s++) {
      re += part[((size_t)q * S + s) * per + e];
      im += part[((size_t)q * S + s) * per + n2 + e];
    }
    return z2{(double)re, (double)im};
  };
  if (g < nles) {
    const int t = g;
    const z2 lt = item(p);
    // Sigma^<(t,T)_ki = -conj(Sigma^<(T,t)_ik)
    stz(Sles + (rowT + t) * n2 + k * n + i, z2{-lt.x, lt.y});
    const int pg = g + nles - p0;  // the ">" partner's place in this chunk
    if (pg < P) {
      const z2 gt = item(pg);
      stz(Sret + (rowT + t) * n2 + e, z2{gt.x - lt.x, gt.y - lt.y});
    } else {
      stz(Sret + (rowT + t) * n2 + e, z2{-lt.x, -lt.y});
    }
  } else if (g < 2 * nles) {
    const int t = g - nles;
    if (t >= p0) return;  // the "<" partner is in this chunk and wrote Sigma^R
    const z2 gt = item(p), r = ldz(Sret + (rowT + t) * n2 + e);
    stz(Sret + (rowT + t) * n2 + e, z2{gt.x + r.x, gt.y + r.y});
  } else {
    const int tau = g - 2 * nles;
    Stv[((size_t)T_ * np + tau) * n2 + e] = item(p);
  }
}

// kern_sum_scatter into the context's Sigma storage (complex FP64, or FP32)
template <typename T>
void sum_scatter(Ctx &c, const T *part, int S, int p0, int P, int T_, hipStream_t s) {
  const long long tot = (long long)P * c.es;
  const dim3 grid((unsigned)((tot + 255) / 256));
  if (c.s32)
    hipLaunchKernelGGL(kern_sum_scatter<T>, grid, dim3(256), 0, s, part, S, p0, P, T_, c.nao, c.np,
                       c.S.ret.as<float2>(), c.S.les.as<float2>(), c.S.tv.as<z2>());
  else
    hipLaunchKernelGGL(kern_sum_scatter<T>, grid, dim3(256), 0, s, part, S, p0, P, T_, c.nao, c.np,
                       c.S.ret.as<z2>(), c.S.les.as<z2>(), c.S.tv.as<z2>());
}

template <typename T>
struct Bufs {
  T *A, *B, *Bt, *W, *X, *H, *Pi, *Ep;
  T *XF;                // folded exchange operand (split exchange, generic path)
  uint16_t *S1h, *S1l;  // split-bf16 operands (bf16x3 path)
  float *Pf;            // [Wr; Wi][Zr; Zi]^T per item (bf16x3 path)
  bool tune;            // let the split-bf16 GEMMs time their tile variants
};

// aux-index splits of the E reduction (K = naux nao is long, M = N = nao small)
inline int e_splits(int a) {
  for (int S : {16, 8, 4, 2})
    if (a % S == 0 && a / S >= 8) return S;
  return 1;
}

inline bool fast_bf16(const Ctx &c, Prec prec) { return prec == PREC_BF16X3 && c.uaug; }

// One chunk of P items starting at item p0.
template <typename T>
void run_chunk(Ctx &c, Prec prec, int T_, int p0, int P, const Bufs<T> &bf, const T *Vq, const T *U) {
  const int n = c.nao, a = c.naux;
  const long long n2 = (long long)n * n, an2 = (long long)a * n2;
  hipStream_t s = c.stream;
  if (c.fast64 && prec == PREC_BF16X3) {
    // nao = 64: every product of the chain on split-bf16 MFMA, each operand
    // split once, intermediates handed over in the layout the next step reads.
    // kern_items left A^T in A, B in B and B^T in Bt.
    constexpr int G = 16;  // aux indices per workgroup
    const SymGeom sg = sym_geom(n, a);
    const bool sym = c.symx;
    const long long ka = sym ? sg.kx : n2 + a;  // row stride of the exchange operand
    const int M = 2 * a * P, NN = (int)n2;
    float *W = (float *)bf.W, *H = (float *)bf.H;
    uint16_t *wth = reinterpret_cast<uint16_t *>(H), *wtl = wth + (size_t)P * 2 * an2;  // Wt lives in H until H is formed (for the whole chunk when split)
    c64_w(c.Vh.as<uint16_t>(), c.Vl.as<uint16_t>(), (const float *)bf.A, P, a, G, sym ? nullptr : W, wth, wtl, s);
    c64_z(c.Vh.as<uint16_t>(), c.Vl.as<uint16_t>(), (const float *)bf.Bt, P, a, G, bf.S1h, bf.S1l, s);
    // Pf_p = [Wt_r; Wt_i] [Zt_r; Zt_i]^T (K index (k,i) in both)
    gemm_nt_bf16x3(NtOp{wth, wtl, n2, 2 * an2}, NtOp{bf.S1h, bf.S1l, n2, 2 * an2}, bf.Pf, 2LL * a, 4LL * a * a, 2 * a,
                   2 * a, NN, P, 1.0f, 0, s, bf.tune);
    pi_combine(bf.Pf, a, P, bf.S1h, bf.S1l, ka, sym ? sg.ns : NN, s);
    check_launch("sigma pi combine");
    if (sym) {
      // the exchange in its symmetric and antisymmetric parts (half the flops
      // of the full contraction); E reads the two outputs as they are
      c64_c((const float *)bf.B, wth, wtl, P, a, G, bf.S1h, bf.S1l, ka, s, c.sym_pairs.as<uint16_t>(), sg.ks);
      float *Y = (float *)bf.X;
      c.mark(0);
      gemm_nt_bf16x3(NtOp{bf.S1h, bf.S1l, ka, 0}, NtOp{c.Ush.as<uint16_t>(), c.Usl.as<uint16_t>(), sg.ks, 0}, Y,
                     sg.ldh, 0, M, sg.nsp, (int)sg.ks, 1, 1.0f, 0, s, bf.tune);
      gemm_nt_bf16x3(NtOp{bf.S1h + sg.ks, bf.S1l + sg.ks, ka, 0},
                     NtOp{c.Uxh.as<uint16_t>(), c.Uxl.as<uint16_t>(), sg.ka, 0}, Y + sg.nsp, sg.ldh, 0, M, sg.nap,
                     (int)sg.ka, 1, 1.0f, 0, s, bf.tune);
      c.mark(1);
      if (c.prof) c.prof_flops += 2.0 * M * ((double)sg.ns * (sg.ns + a) + (double)sg.na * sg.na);
    } else {
      c64_c((const float *)bf.B, wth, wtl, P, a, G, bf.S1h, bf.S1l, ka, s);
      // Ht^a = ([C | Pi] Baug^T) with Baug rows ordered (j,k): H^a transposed
      c.mark(0);
      gemm_nt_bf16x3(NtOp{bf.S1h, bf.S1l, ka, 0}, NtOp{c.Uah.as<uint16_t>(), c.Ual.as<uint16_t>(), ka, 0}, H, n2, 0,
                     M, NN, (int)ka, 1, 1.0f, 0, s, bf.tune);
      c.mark(1);
      if (c.prof) c.prof_flops += 2.0 * M * (double)NN * (double)ka;
    }
    const int S = e_splits(a);
    if (sym)
      c64_e(nullptr, nullptr, P, a, S, (float *)bf.Ep, s, (const float *)bf.X, sg.ldh, sg.nsp,
            c.sym_pairs.as<uint16_t>(), wth, wtl);
    else c64_e(W, H, P, a, S, (float *)bf.Ep, s);
    sum_scatter(c, bf.Ep, S, p0, P, T_, s);
    check_launch("sigma sum scatter");
    return;
  }
  // W_p = Vstack A_p : (a n x n)(n x n), Vstack[(a,i)][l] = B[a][i][l]
  {
    GemmDesc d;
    d.M = a * n; d.N = n; d.K = n; d.batch = P; d.ca = false; d.cb = true;
    d.A.ptr = Vq; d.A.ms = n; d.A.ks_lo = 1;
    d.B.ptr = bf.A; d.B.ms = 1; d.B.ks_lo = n; d.B.ps = n2; d.B.bs_lo = 2 * n2;
    d.C.ptr = bf.W; d.C.rs = n; d.C.cs = 1; d.C.ps = an2; d.C.bs_lo = 2 * an2;
    gemm(d, prec, s);
  }
  // Z^b = B_p^T (V^b)^T  (into X):  Z^b_ik = sum_l B_li V^b_kl
  {
    GemmDesc d;
    d.M = n; d.N = n; d.K = n; d.batch = P * a; d.ca = true; d.cb = false;
    d.A.ptr = bf.B; d.A.ms = 1; d.A.ks_lo = n; d.A.ps = n2; d.A.bdiv = a; d.A.bs_hi = 2 * n2; d.A.bs_lo = 0;
    d.B.ptr = Vq; d.B.ms = n; d.B.ks_lo = 1; d.B.bdiv = a; d.B.bs_hi = 0; d.B.bs_lo = n2;
    d.C.ptr = bf.X; d.C.rs = n; d.C.cs = 1; d.C.ps = an2; d.C.bdiv = a; d.C.bs_hi = 2 * an2; d.C.bs_lo = n2;
    gemm(d, prec, s);
  }
  const bool fast = fast_bf16(c, prec);
  const int M = 2 * a * P, NN = (int)n2;
  if (fast) {
    // split-bf16 operands: W into the (still unused) H buffer, Z into S1
    uint16_t *wh = reinterpret_cast<uint16_t *>(bf.H), *wl = wh + (size_t)P * 2 * an2;
    split_bf16((const float *)bf.W, n2, wh, wl, n2, (long long)M, NN, s);
    split_bf16((const float *)bf.X, n2, bf.S1h, bf.S1l, n2, (long long)M, NN, s);
    // Pf_p = [Wr; Wi] [Zr; Zi]^T: M = N = 2a, K = n^2, one launch for the chunk
    gemm_nt_bf16x3(NtOp{wh, wl, n2, 2 * an2}, NtOp{bf.S1h, bf.S1l, n2, 2 * an2}, bf.Pf, 2LL * a, 4LL * a * a, 2 * a,
                   2 * a, NN, P, 1.0f, 0, s, bf.tune);
  } else {
    // Pi_p[a][b] = sum_(ik) W^a_ik Z^b_ik
    GemmDesc d;
    d.M = a; d.N = a; d.K = (int)n2; d.batch = P; d.ca = true; d.cb = true;
    d.A.ptr = bf.W; d.A.ms = n2; d.A.ks_lo = 1; d.A.ps = an2; d.A.bs_lo = 2 * an2;
    d.B.ptr = bf.X; d.B.ms = n2; d.B.ks_lo = 1; d.B.ps = an2; d.B.bs_lo = 2 * an2;
    d.C.ptr = bf.Pi; d.C.rs = a; d.C.cs = 1; d.C.ps = (long long)a * a; d.C.bs_lo = 2LL * a * a;
    gemm(d, prec, s);
  }
  // C^a = B_p W^a  (into X)
  {
    GemmDesc d;
    d.M = n; d.N = n; d.K = n; d.batch = P * a; d.ca = true; d.cb = true;
    d.A.ptr = bf.B; d.A.ms = n; d.A.ks_lo = 1; d.A.ps = n2; d.A.bdiv = a; d.A.bs_hi = 2 * n2; d.A.bs_lo = 0;
    d.B.ptr = bf.W; d.B.ms = 1; d.B.ks_lo = n; d.B.ps = an2; d.B.bdiv = a; d.B.bs_hi = 2 * an2; d.B.bs_lo = n2;
    d.C.ptr = bf.X; d.C.rs = n; d.C.cs = 1; d.C.ps = an2; d.C.bdiv = a; d.C.bs_hi = 2 * an2; d.C.bs_lo = n2;
    gemm(d, prec, s);
  }
  if (fast) {
    // H = [C | Pi] [-U'^T | 2 Vmat^T]^T = -C U' + 2 Pi Vmat: the exchange and the
    // bubble's Pi V in one GEMM, rows (item, plane, a), K = n^2 + a
    const long long ka = n2 + a;
    pi_combine(bf.Pf, a, P, bf.S1h, bf.S1l, ka, NN, s);
    check_launch("sigma pi combine");
    split_bf16((const float *)bf.X, n2, bf.S1h, bf.S1l, ka, (long long)M, NN, s);
    c.mark(0);
    gemm_nt_bf16x3(NtOp{bf.S1h, bf.S1l, ka, 0}, NtOp{c.Uah.as<uint16_t>(), c.Ual.as<uint16_t>(), ka, 0},
                   (float *)bf.H, n2, 0, M, NN, (int)ka, 1, 1.0f, 0, s, bf.tune);
    c.mark(1);
    if (c.prof) c.prof_flops += 2.0 * M * (double)NN * (double)ka;
  } else if (c.symg) {
    // the exchange split by the pair symmetry of U' (kern_build_Usym_gen):
    // fold C and Pi into the rows [xs | Pi | xa], two GEMMs into Y = [Cs | Ca]
    // (in X, C being folded), then H^a unfolded from Y
    const int ns = n * (n + 1) / 2, na = n * (n - 1) / 2;
    const long long kx = ns + a + na, ly = ns + na;
    const long long tot = (long long)M * kx;
    hipLaunchKernelGGL(kern_fold_pairs<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, bf.X, bf.Pi,
                       c.sym_pairs.as<uint16_t>(), n, a, (long long)M, bf.XF);
    check_launch("sigma fold pairs");
    const T *Usg = sizeof(T) == 8 ? (const T *)c.Usg64.p : (const T *)c.Usg32.p;
    c.mark(0);
    {
      GemmDesc d;
      d.M = M; d.N = ns; d.K = ns + a; d.batch = 1;
      d.A.ptr = bf.XF; d.A.ms = kx; d.A.ks_lo = 1;
      d.B.ptr = Usg; d.B.ms = 1; d.B.ks_lo = ns;
      d.C.ptr = bf.X; d.C.rs = ly; d.C.cs = 1;
      gemm(d, prec, s);
    }
    {
      GemmDesc d;
      d.M = M; d.N = na; d.K = na; d.batch = 1;
      d.A.ptr = bf.XF + ns + a; d.A.ms = kx; d.A.ks_lo = 1;
      d.B.ptr = Usg + (size_t)(ns + a) * ns; d.B.ms = 1; d.B.ks_lo = na;
      d.C.ptr = bf.X + ns; d.C.rs = ly; d.C.cs = 1;
      gemm(d, prec, s);
    }
    c.mark(1);
    if (c.prof) c.prof_flops += 2.0 * M * ((double)ns * (ns + a) + (double)na * na);
    const long long toth = (long long)M * n2;
    hipLaunchKernelGGL(kern_unfold_pairs<T>, dim3((unsigned)((toth + 255) / 256)), dim3(256), 0, s, bf.X, n,
                       (long long)M, bf.H);
    check_launch("sigma unfold pairs");
  } else {
    // H = -C U'   (rows (p, plane, a) of X, all at once; the n^4 naux GEMM)
    {
      GemmDesc d;
      d.M = M; d.N = NN; d.K = NN; d.batch = 1;
      d.A.ptr = bf.X; d.A.ms = n2; d.A.ks_lo = 1;
      d.B.ptr = U; d.B.ms = 1; d.B.ks_lo = n2;
      d.C.ptr = bf.H; d.C.rs = n2; d.C.cs = 1;
      d.alpha = -1.0;
      c.mark(0);
      gemm(d, prec, s);
      c.mark(1);
      if (c.prof) c.prof_flops += 2.0 * M * (double)NN * NN;
    }
    // H += 2 Pi Vmat   (Vmat[b][(k,j)] = B[b][k][j])
    {
      GemmDesc d;
      d.M = a; d.N = (int)n2; d.K = a; d.batch = P; d.ca = true; d.cb = false;
      d.A.ptr = bf.Pi; d.A.ms = a; d.A.ks_lo = 1; d.A.ps = (long long)a * a; d.A.bs_lo = 2LL * a * a;
      d.B.ptr = Vq; d.B.ms = 1; d.B.ks_lo = n2;
      d.C.ptr = bf.H; d.C.rs = n2; d.C.cs = 1; d.C.ps = an2; d.C.bs_lo = 2 * an2;
      d.alpha = 2.0; d.beta = 1;
      gemm(d, prec, s);
    }
  }
  // E_p = sum_a W^a H^a  (K index (a,k): W offset a*n2 + k), split over S
  // ranges of a so that the launch has P*S workgroups, then summed
  {
    const int S = e_splits(a), as = a / S;
    GemmDesc d;
    d.M = n; d.N = n; d.K = as * n; d.batch = P * S; d.ca = true; d.cb = true;
    d.A.ptr = bf.W; d.A.ms = n; d.A.kdiv = n; d.A.ks_hi = n2; d.A.ks_lo = 1; d.A.ps = an2;
    d.A.bdiv = S; d.A.bs_hi = 2 * an2; d.A.bs_lo = (long long)as * n2;
    d.B.ptr = bf.H; d.B.ms = 1; d.B.ks_lo = n; d.B.ps = an2;
    d.B.bdiv = S; d.B.bs_hi = 2 * an2; d.B.bs_lo = (long long)as * n2;
    d.C.ptr = bf.Ep; d.C.rs = n; d.C.cs = 1; d.C.ps = n2; d.C.bs_lo = 2 * n2;
    gemm(d, prec, s);
    long long tot = (long long)P * n2;
    sum_scatter(c, bf.Ep, S, p0, P, T_, s);
    check_launch("sigma sum scatter");
  }
}

// The scratch of one chunk of P items, carved from one buffer at base (sizes
// only when base is null): items A, B (and B^T on the nao = 64 path); W
// (not kept on the split path: E reads W^a from its bf16 split in H); X (the
// generic path's Z and C; on the split path the two exchange outputs); H;
// the E partial sums; and on the bf16x3 path the split operand S1 and the Pi
// product Pf, on the others Pi.  Every piece starts on a 256-byte boundary.
template <typename T>
size_t sigma_carve(const Ctx &c, Prec prec, int P, char *base, Bufs<T> *bf) {
  const int n = c.nao, a = c.naux;
  const size_t n2 = (size_t)n * n, an2 = (size_t)a * n2, Pz = (size_t)P;
  const bool fast = fast_bf16(c, prec), f64w = fast && c.fast64, sym = f64w && c.symx;
  const SymGeom sg = sym_geom(n, a);
  const size_t ka = sym ? (size_t)sg.kx : n2 + a;
  const int S = e_splits(a);
  size_t off = 0;
  auto take = [&](size_t bytes) -> char * {
    char *q = base ? base + off : nullptr;
    off += (bytes + 255) / 256 * 256;
    return bytes ? q : nullptr;
  };
  Bufs<T> b{};
  b.A = (T *)take(Pz * 2 * n2 * sizeof(T));
  b.B = (T *)take(Pz * 2 * n2 * sizeof(T));
  b.Bt = f64w ? (T *)take(Pz * 2 * n2 * sizeof(T)) : nullptr;
  b.W = sym ? nullptr : (T *)take(Pz * 2 * an2 * sizeof(T));
  if (sym) b.X = (T *)take(Pz * 2 * a * (size_t)sg.ldh * sizeof(float));
  else if (!f64w) b.X = (T *)take(Pz * 2 * an2 * sizeof(T));
  b.H = (T *)take(Pz * 2 * an2 * sizeof(T));
  b.Ep = (T *)take(Pz * S * 2 * n2 * sizeof(T));
  if (fast) {
    b.S1h = (uint16_t *)take(Pz * 2 * a * ka * sizeof(uint16_t));
    b.S1l = (uint16_t *)take(Pz * 2 * a * ka * sizeof(uint16_t));
    b.Pf = (float *)take(Pz * 4 * a * a * sizeof(float));
  } else {
    b.Pi = (T *)take(Pz * 2 * a * a * sizeof(T));
    if (c.symg) b.XF = (T *)take(Pz * 2 * a * (n2 + a) * sizeof(T));  // ns + a + na = n^2 + a columns
  }
  b.tune = true;
  if (bf) *bf = b;
  return off;
}

// Sigma of slice tstp in chunks of P items, with the scratch at base (sized
// by sigma_carve for P)
template <typename T>
void sigma_chunks(Ctx &c, int tstp, Prec prec, int P, char *base, bool tune) {
  const int n = c.nao, np = c.np;
  const long long n2 = (long long)n * n;
  const int Ptot = 2 * (tstp + 1) + np;
  const bool f64w = fast_bf16(c, prec) && c.fast64;
  Bufs<T> bf;
  sigma_carve<T>(c, prec, P, base, &bf);
  bf.tune = tune;
  const T *Vq = sizeof(T) == 8 ? (const T *)c.Bq.p : (const T *)c.Bq32.p;
  const T *U = sizeof(T) == 8 ? (const T *)c.U64.p : (const T *)c.U32.p;
  for (int p0 = 0; p0 < Ptot; p0 += P) {
    int Pc = std::min(P, Ptot - p0);
    long long tot = (long long)Pc * n2;
    hipLaunchKernelGGL(kern_items<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c.stream, c.G.ret.as<z2>(),
                       c.G.les.as<z2>(), c.G.tv.as<z2>(), tstp, g_row(c, tstp), n, np, c.sig, p0, Pc, bf.A, bf.B,
                       f64w ? 1 : 0,
                       bf.Bt);
    check_launch("sigma items");
    run_chunk<T>(c, prec, tstp, p0, Pc, bf, Vq, U);
  }
}

template <typename T>
void sigma_typed(Ctx &c, int tstp, Prec prec) {
  const int Ptot = 2 * (tstp + 1) + c.np;
  const size_t per_item = sigma_carve<T>(c, prec, 1, nullptr, nullptr);
  const size_t budget = (size_t)6 << 30;
  int P = c.sigma_chunk > 0 ? c.sigma_chunk : (int)std::max<size_t>(1, budget / per_item);
  if (c.sigma_chunk <= 0 && P >= 16) {
    // a multiple of 8 items: the exchange GEMM of a full chunk then has
    // naux*P/8 tiles of 256 x 256 (at nao = 64), whole rounds of 256 CUs;
    // rounded up when that stays within 5% of the scratch budget
    const int up = (P + 7) / 8 * 8;
    P = (double)up * per_item <= 1.05 * budget ? up : P / 8 * 8;
  }
  // at least 464 items a chunk where that takes at most 16 GiB: at nao = 64
  // (26 MB an item) chunks of 464 ran the Sigma evaluation 3-4% faster than
  // chunks of 232, 352 or 712 (1066: 5% slower); systems of smaller nao reach
  // larger chunks within the budget anyway
  constexpr int kPmin = 464;
  if (c.sigma_chunk <= 0 && P < kPmin && (double)kPmin * per_item <= (double)((size_t)16 << 30)) P = kPmin;
  P = std::min(P, Ptot);
  c.sg_arena.ensure(sigma_carve<T>(c, prec, P, nullptr, nullptr));
  sigma_chunks<T>(c, tstp, prec, P, c.sg_arena.as<char>(), true);
}

// ---- the same computation with no device allocation: the chunk scratch is
// this file's static device array (kScratchBytes per GPU, held until the
// process ends), shared by every call on that GPU
constexpr size_t kScratchBytes = (size_t)256 << 20;
static __device__ __align__(256) char scratch[kScratchBytes];
std::mutex g_scratch_mu;
std::map<int, char *> g_scratch_ptr;  // device -> address of that device's copy of scratch

}  // namespace

void sigma_scratch_init(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr.count(device)) return;
  int cur = -1;
  NED_HIP(hipGetDevice(&cur));
  require(cur == device, "sigma scratch: the device must be current");
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
  g_scratch_ptr[device] = (char *)ptr;
}

size_t sigma_scratch_bytes() { return kScratchBytes; }

template <typename T>
static void sigma_noalloc_typed(Ctx &c, int tstp, Prec prec) {
  char *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(c.device);
    if (it == g_scratch_ptr.end()) throw std::runtime_error("state: sigma scratch not initialised for this device");
    base = it->second;
  }
  const int Ptot = 2 * (tstp + 1) + c.np;
  const size_t per_item = sigma_carve<T>(c, prec, 1, nullptr, nullptr);
  if (per_item > kScratchBytes)
    throw std::invalid_argument("sigma (no allocation): one time point needs " + std::to_string(per_item) +
                                " bytes of scratch, more than the " + std::to_string(kScratchBytes) + " available");
  int P = std::min<long long>(Ptot, (long long)(kScratchBytes / per_item));
  while (P > 1 && sigma_carve<T>(c, prec, P, nullptr, nullptr) > kScratchBytes) P--;
  // no event timing (hipEventCreate) and no tile autotuning on this path
  const bool prof = c.prof;
  c.prof = false;
  try {
    sigma_chunks<T>(c, tstp, prec, P, base, false);
  } catch (...) {
    c.prof = prof;
    throw;
  }
  c.prof = prof;
}

void sigma_gf2_noalloc(Ctx &c, int tstp, Prec prec) {
  require(c.have_coulomb, "sigma: Coulomb tensor not set");
  require(tstp >= 0 && tstp <= c.nt, "sigma: tstp out of range");
  g_need(c, tstp, tstp, "sigma");
  if (prec == PREC_F64) sigma_noalloc_typed<double>(c, tstp, prec);
  else sigma_noalloc_typed<float>(c, tstp, prec);
}

void build_coulomb(Ctx &c, const double *Bq_host) {
  const int n = c.nao, a = c.naux;
  const size_t n2 = (size_t)n * n;
  for (int Q = 0; Q < a; Q++)
    for (int i = 0; i < n; i++)
      for (int j = 0; j < i; j++) {
        double x = Bq_host[Q * n2 + i * n + j], y = Bq_host[Q * n2 + j * n + i];
        if (std::fabs(x - y) > 1e-12 * (1.0 + std::fabs(x)))
          throw std::invalid_argument("Coulomb factor B[Q,i,j] must be symmetric in (i,j)");
      }
  for (DevBuf *b : {&c.U4_64, &c.U4_32, &c.Wx64, &c.Wx32}) b->release();  // the full solver's, rebuilt on use
  c.Bq.ensure(a * n2 * sizeof(double));
  c.Bq32.ensure(a * n2 * sizeof(float));
  c.U64.ensure(n2 * n2 * sizeof(double));
  c.U32.ensure(n2 * n2 * sizeof(float));
  NED_HIP(hipMemcpyAsync(c.Bq.p, Bq_host, a * n2 * sizeof(double), hipMemcpyHostToDevice, c.stream));
  long long nb = (long long)a * n2;
  hipLaunchKernelGGL(kern_d2f, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, c.stream, c.Bq.as<double>(),
                     c.Bq32.as<float>(), nb);
  check_launch("coulomb d2f");
  long long nu = (long long)n2 * n2;
  hipLaunchKernelGGL(kern_build_U, dim3((unsigned)((nu + 255) / 256)), dim3(256), 0, c.stream, c.Bq.as<double>(), n, a,
                     c.U64.as<double>(), c.U32.as<float>());
  check_launch("coulomb U");
  c.uaug = n2 % 128 == 0 && (2 * a) % 128 == 0 && a % 64 == 0;
  c.fast64 = c.uaug && n == 64;
  // at nao = 64 the exchange runs split by the symmetry of U' (NED_SYMX=0
  // keeps the full contraction)
  const char *sx = std::getenv("NED_SYMX");
  c.symx = c.fast64 && !(sx && std::atoi(sx) == 0);
  // the full contraction's operand, not read by the split exchange
  if (c.symx) {
    c.Uah.release();
    c.Ual.release();
  }
  if (c.uaug && !c.symx) {
    const size_t ka = n2 + a;
    c.Uah.ensure(n2 * ka * sizeof(uint16_t));
    c.Ual.ensure(n2 * ka * sizeof(uint16_t));
    long long nt_ = (long long)(n2 * ka);
    hipLaunchKernelGGL(kern_build_Uaug, dim3((unsigned)((nt_ + 255) / 256)), dim3(256), 0, c.stream, c.U32.as<float>(),
                       c.Bq32.as<float>(), n, a, c.fast64 ? 1 : 0, c.Uah.as<uint16_t>(), c.Ual.as<uint16_t>());
    check_launch("coulomb Baug");
  }
  if (c.fast64) {
    c.Vh.ensure(a * n2 * sizeof(uint16_t));
    c.Vl.ensure(a * n2 * sizeof(uint16_t));
    split_bf16(c.Bq32.as<float>(), (long long)n2, c.Vh.as<uint16_t>(), c.Vl.as<uint16_t>(), (long long)n2, a, (int)n2,
               c.stream);
  }
  // the split on the FP64 / FP32 path (and bf16x3 without the split-bf16
  // GEMMs) from nao = 8 on
  c.symg = !(sx && std::atoi(sx) == 0) && n >= 8;
  if (c.symx || c.symg) {
    const int npair = n * (n + 1) / 2 + n * (n - 1) / 2;
    c.sym_pairs.ensure((size_t)npair * sizeof(uint16_t));
    hipLaunchKernelGGL(kern_sym_pairs, dim3((npair + 255) / 256), dim3(256), 0, c.stream, n,
                       c.sym_pairs.as<uint16_t>());
    check_launch("coulomb pairs");
  }
  if (c.symg) {
    const long long ns = n * (n + 1) / 2, na = n * (n - 1) / 2, tot = (ns + a) * ns + na * na;
    c.Usg64.ensure((size_t)tot * sizeof(double));
    c.Usg32.ensure((size_t)tot * sizeof(float));
    hipLaunchKernelGGL(kern_build_Usym_gen<double>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c.stream,
                       c.U64.as<double>(), c.Bq.as<double>(), c.sym_pairs.as<uint16_t>(), n, a, c.Usg64.as<double>());
    hipLaunchKernelGGL(kern_build_Usym_gen<float>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c.stream,
                       c.U64.as<double>(), c.Bq.as<double>(), c.sym_pairs.as<uint16_t>(), n, a, c.Usg32.as<float>());
    check_launch("coulomb split U (FP64 / FP32)");
  } else {
    c.Usg64.release();
    c.Usg32.release();
  }
  if (c.symx) {
    const SymGeom sg = sym_geom(n, a);
    // kern_c64_c zeroes 32 padding columns after each block (naux % 64 == 0)
    require(sg.ks - sg.ns - a == 32 && sg.ka - sg.na == 32, "split exchange: unexpected padding");
    c.Ush.ensure((size_t)sg.nsp * sg.ks * sizeof(uint16_t));
    c.Usl.ensure((size_t)sg.nsp * sg.ks * sizeof(uint16_t));
    c.Uxh.ensure((size_t)sg.nap * sg.ka * sizeof(uint16_t));
    c.Uxl.ensure((size_t)sg.nap * sg.ka * sizeof(uint16_t));
    long long ns_ = (long long)sg.nsp * sg.ks, na_ = (long long)sg.nap * sg.ka;
    hipLaunchKernelGGL(kern_build_Usym, dim3((unsigned)((ns_ + 255) / 256)), dim3(256), 0, c.stream,
                       c.U64.as<double>(), c.Bq.as<double>(), c.sym_pairs.as<uint16_t>(), n, a, 0, sg.nsp,
                       (int)sg.ks, c.Ush.as<uint16_t>(), c.Usl.as<uint16_t>());
    hipLaunchKernelGGL(kern_build_Usym, dim3((unsigned)((na_ + 255) / 256)), dim3(256), 0, c.stream,
                       c.U64.as<double>(), c.Bq.as<double>(), c.sym_pairs.as<uint16_t>(), n, a, 1, sg.nap,
                       (int)sg.ka, c.Uxh.as<uint16_t>(), c.Uxl.as<uint16_t>());
    check_launch("coulomb split U");
  }
  NED_HIP(hipStreamSynchronize(c.stream));
  c.have_coulomb = true;
}

void sigma_gf2(Ctx &c, int tstp, Prec prec) {
  require(c.have_coulomb, "sigma: Coulomb tensor not set");
  require(tstp >= 0 && tstp <= c.nt, "sigma: tstp out of range");
  g_need(c, tstp, tstp, "sigma");
  c.mark(2);
  if (prec == PREC_F64) sigma_typed<double>(c, tstp, prec);
  else sigma_typed<float>(c, tstp, prec);
  c.mark(3);
}

// ---- the full (non-decomposed) second-Born self-energy, molGF2Solver::solve
// (src/mol2bSolver_impl.h:5-120): the same functional F[A,B] of every item,
// contracted through the four-index U instead of the Cholesky vectors,
//   T1[i,k,n,p] = sum_l A_lk U_{il,np}        (per item and i: n x n^2 x n)
//   T2[i,k,n,q] = sum_p T1[i,k,n,p] A_pq      (n^3 x n x n)
//   T3[i,k,q,m] = sum_n T2[i,k,n,q] B_mn      (per (i,k): n x n x n)
//   Sigma_ij    = sum_{kqm} T3[i,k,q,m] Wx[(k,q,m)][j],  Wx = 2 U_{kj,qm} - U_{qj,km}
// (the reference's Uexch, src/mol2bSolver.h:44-48), 24 n^5 flop per item.

namespace {
// The two four-index operands of the full solver from U' as built by
// kern_build_U, U'[(a,b)][(c,d)] = sum_Q B[Q][c][a] B[Q][b][d] = U_{ca,bd}:
//   U4[(i,l)][(n,p)]  = U_{il,np} = U'[(l,n)][(i,p)]
//   Wx[(k,q,m)][j]    = 2 U_{kj,qm} - U_{qj,km} = 2 U'[(j,q)][(k,m)] - U'[(j,k)][(q,m)]
template <typename T>
__global__ void kern_build_full_ops(const double *U, int n, T *U4, T *Wx) {
  const long long n2 = (long long)n * n, n4 = n2 * n2;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n4) return;
  {
    const int p = (int)(idx % n), nn = (int)((idx / n) % n), l = (int)((idx / n2) % n), i = (int)(idx / (n2 * n));
    U4[idx] = (T)U[(l * n + nn) * n2 + i * n + p];
  }
  const int j = (int)(idx % n);
  const long long c = idx / n;
  const int m = (int)(c % n), q = (int)((c / n) % n), k = (int)(c / n2);
  Wx[idx] = (T)(2.0 * U[(j * n + q) * n2 + k * n + m] - U[(j * n + k) * n2 + q * n + m]);
}

template <typename T>
void sigma_full_typed(Ctx &c, int tstp, Prec prec) {
  const int n = c.nao, np = c.np;
  const long long n2 = (long long)n * n, n3 = n2 * n, n4 = n2 * n2;
  const int Ptot = 2 * (tstp + 1) + np;
  const int S = n;  // K splits of the last product (K = n^3)
  // per item: A, B, T1 (also T3), T2 (planar complex) and the S partials of Sigma
  const size_t per_item = ((size_t)4 * n2 + 4 * n4 + (size_t)S * 2 * n2) * sizeof(T) + 4 * 256;
  const size_t budget = (size_t)6 << 30;
  int P = c.sigma_chunk > 0 ? c.sigma_chunk : (int)std::max<size_t>(1, budget / per_item);
  P = std::min(P, Ptot);
  const size_t bA = (size_t)P * 2 * n2 * sizeof(T), bT = (size_t)P * 2 * n4 * sizeof(T),
               bE = (size_t)P * S * 2 * n2 * sizeof(T);
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  c.sg_arena.ensure(2 * up(bA) + 2 * up(bT) + up(bE));
  char *base = c.sg_arena.as<char>();
  T *A = (T *)base, *B = (T *)(base + up(bA)), *T1 = (T *)(base + 2 * up(bA)),
    *T2 = (T *)(base + 2 * up(bA) + up(bT)), *E = (T *)(base + 2 * up(bA) + 2 * up(bT));
  DevBuf &Wb = sizeof(T) == 8 ? c.Wx64 : c.Wx32;
  DevBuf &Ub = sizeof(T) == 8 ? c.U4_64 : c.U4_32;
  if (!Wb.p || !Ub.p) {
    Wb.ensure((size_t)n4 * sizeof(T));
    Ub.ensure((size_t)n4 * sizeof(T));
    hipLaunchKernelGGL(kern_build_full_ops<T>, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, c.stream,
                       c.U64.as<double>(), n, Ub.as<T>(), Wb.as<T>());
    check_launch("sigma full operands");
  }
  const T *U = Ub.as<T>();
  hipStream_t s = c.stream;
  for (int p0 = 0; p0 < Ptot; p0 += P) {
    const int Pc = std::min(P, Ptot - p0);
    const long long Pn4 = (long long)P * n4;
    long long tot = (long long)Pc * n2;
    hipLaunchKernelGGL(kern_items<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, c.G.ret.as<z2>(),
                       c.G.les.as<z2>(), c.G.tv.as<z2>(), tstp, g_row(c, tstp), n, np, c.sig, p0, Pc, A, B, 0,
                       (T *)nullptr);
    check_launch("sigma full items");
    {  // T1[p][i][k][(n,p')] = sum_l A_lk U[(i,l)][(n,p')]
      GemmDesc d;
      d.M = n; d.N = (int)n2; d.K = n; d.batch = Pc * n; d.ca = true; d.cb = false;
      d.A.ptr = A; d.A.ms = 1; d.A.ks_lo = n; d.A.ps = n2; d.A.bdiv = n; d.A.bs_hi = 2 * n2; d.A.bs_lo = 0;
      d.B.ptr = U; d.B.ms = 1; d.B.ks_lo = n2; d.B.bdiv = n; d.B.bs_hi = 0; d.B.bs_lo = n3;
      d.C.ptr = T1; d.C.rs = n2; d.C.cs = 1; d.C.ps = Pn4; d.C.bdiv = n; d.C.bs_hi = n4; d.C.bs_lo = n3;
      gemm(d, prec, s);
    }
    {  // T2[p][(i,k,n)][q] = sum_p' T1[p][(i,k,n)][p'] A_p'q
      GemmDesc d;
      d.M = (int)n3; d.N = n; d.K = n; d.batch = Pc; d.ca = true; d.cb = true;
      d.A.ptr = T1; d.A.ms = n; d.A.ks_lo = 1; d.A.ps = Pn4; d.A.bs_lo = n4;
      d.B.ptr = A; d.B.ms = 1; d.B.ks_lo = n; d.B.ps = n2; d.B.bs_lo = 2 * n2;
      d.C.ptr = T2; d.C.rs = n; d.C.cs = 1; d.C.ps = Pn4; d.C.bs_lo = n4;
      gemm(d, prec, s);
    }
    {  // T3[p][(i,k)][q][m] = sum_n T2[p][(i,k)][n][q] B_mn   (into T1)
      GemmDesc d;
      d.M = n; d.N = n; d.K = n; d.batch = Pc * (int)n2; d.ca = true; d.cb = true;
      d.A.ptr = T2; d.A.ms = 1; d.A.ks_lo = n; d.A.ps = Pn4; d.A.bdiv = (int)n2; d.A.bs_hi = n4; d.A.bs_lo = n2;
      d.B.ptr = B; d.B.ms = n; d.B.ks_lo = 1; d.B.ps = n2; d.B.bdiv = (int)n2; d.B.bs_hi = 2 * n2; d.B.bs_lo = 0;
      d.C.ptr = T1; d.C.rs = n; d.C.cs = 1; d.C.ps = Pn4; d.C.bdiv = (int)n2; d.C.bs_hi = n4; d.C.bs_lo = n2;
      gemm(d, prec, s);
    }
    {  // partial Sigma_ij over K range s: sum_c T3[p][i][c] Wx[c][j], c in [s n^2, (s+1) n^2)
      GemmDesc d;
      d.M = n; d.N = n; d.K = (int)(n3 / S); d.batch = Pc * S; d.ca = true; d.cb = false;
      d.A.ptr = T1; d.A.ms = n3; d.A.ks_lo = 1; d.A.ps = Pn4; d.A.bdiv = S; d.A.bs_hi = n4; d.A.bs_lo = n3 / S;
      d.B.ptr = Wb.p; d.B.ms = 1; d.B.ks_lo = n; d.B.bdiv = S; d.B.bs_hi = 0; d.B.bs_lo = (n3 / S) * n;
      d.C.ptr = E; d.C.rs = n; d.C.cs = 1; d.C.ps = n2; d.C.bdiv = S; d.C.bs_hi = (long long)S * 2 * n2;
      d.C.bs_lo = 2 * n2;
      gemm(d, prec, s);
    }
    tot = (long long)Pc * n2;
    sum_scatter(c, E, S, p0, Pc, tstp, s);
    check_launch("sigma full scatter");
  }
}
}  // namespace

void sigma_gf2_full(Ctx &c, int tstp, Prec prec) {
  require(c.have_coulomb, "sigma: Coulomb tensor not set");
  require(tstp >= 0 && tstp <= c.nt, "sigma: tstp out of range");
  g_need(c, tstp, tstp, "sigma");
  if (prec == PREC_F64) sigma_full_typed<double>(c, tstp, prec);
  else sigma_full_typed<float>(c, tstp, prec);
}

}  // namespace ned
