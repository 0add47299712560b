// Workgroup-level operations on nao x nao complex FP64 blocks (nao <= 64):
// the building block of the Dyson-step history sums.  256 threads form a
// 16 x 16 grid; thread (ty, tx) owns rows ty+16r and columns tx+16c,
// r, c = 0..3, of the output block, accumulated in registers.  Operand
// K-slices of 16 are staged through LDS.
#pragma once
#include "common.h"

namespace ned {

constexpr int BO_T = 256;   // threads per workgroup
constexpr int BO_KT = 16;   // K slice
constexpr int BO_NMAX = 64; // largest supported nao

struct BlkAcc {
  z2 v[4][4];
  __device__ void zero() {
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
      for (int c = 0; c < 4; c++) v[r][c] = {0.0, 0.0};
  }
};

struct BlkLds {
  z2 X[BO_NMAX][BO_KT + 1];
  z2 Y[BO_KT][BO_NMAX + 1];
};

// op: 0 = as stored, 1 = adjoint (conjugate transpose), 2 = transpose
__device__ __forceinline__ z2 blk_get(const z2 *M, int n, int op, int r, int c) {
  if (op == 0) return M[r * n + c];
  z2 v = M[c * n + r];
  return op == 1 ? zconj(v) : v;
}

// acc += coef * op(X) * op(Y)   (all threads of the workgroup call this)
__device__ __forceinline__ void blk_mac(BlkAcc &acc, BlkLds &s, int n, const z2 *X, int opX, const z2 *Y, int opY,
                                        z2 coef) {
  const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
  for (int k0 = 0; k0 < n; k0 += BO_KT) {
    __syncthreads();
    for (int e = tid; e < BO_NMAX * BO_KT; e += BO_T) {
      int r = e / BO_KT, kk = e % BO_KT;
      z2 v = {0.0, 0.0};
      if (r < n && k0 + kk < n) v = zmul(coef, blk_get(X, n, opX, r, k0 + kk));
      s.X[r][kk] = v;
    }
    for (int e = tid; e < BO_KT * BO_NMAX; e += BO_T) {
      int kk = e / BO_NMAX, c = e % BO_NMAX;
      z2 v = {0.0, 0.0};
      if (c < n && k0 + kk < n) v = blk_get(Y, n, opY, k0 + kk, c);
      s.Y[kk][c] = v;
    }
    __syncthreads();
    const int kmax = (n - k0) < BO_KT ? (n - k0) : BO_KT;
    for (int kk = 0; kk < kmax; kk++) {
      z2 a[4], b[4];
#pragma unroll
      for (int r = 0; r < 4; r++) a[r] = s.X[ty + 16 * r][kk];
#pragma unroll
      for (int c = 0; c < 4; c++) b[c] = s.Y[kk][tx + 16 * c];
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
          acc.v[r][c].x = fma(a[r].x, b[c].x, fma(-a[r].y, b[c].y, acc.v[r][c].x));
          acc.v[r][c].y = fma(a[r].x, b[c].y, fma(a[r].y, b[c].x, acc.v[r][c].y));
        }
    }
  }
}

// acc += coef * op(X)   (elementwise, no product)
__device__ __forceinline__ void blk_axpy(BlkAcc &acc, int n, const z2 *X, int opX, z2 coef) {
  const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      int i = ty + 16 * r, j = tx + 16 * c;
      if (i < n && j < n) acc.v[r][c] = zadd(acc.v[r][c], zmul(coef, blk_get(X, n, opX, i, j)));
    }
}

__device__ __forceinline__ void blk_load(BlkAcc &acc, int n, const z2 *X) {
  const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      int i = ty + 16 * r, j = tx + 16 * c;
      acc.v[r][c] = (i < n && j < n) ? X[i * n + j] : z2{0.0, 0.0};
    }
}

__device__ __forceinline__ void blk_store(const BlkAcc &acc, int n, z2 *X) {
  const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      int i = ty + 16 * r, j = tx + 16 * c;
      if (i < n && j < n) X[i * n + j] = acc.v[r][c];
    }
}

__device__ __forceinline__ void blk_scale(BlkAcc &acc, z2 f) {
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc.v[r][c] = zmul(acc.v[r][c], f);
}

// squared Frobenius distance between acc and X, reduced over the workgroup
// (result valid in thread 0)
__device__ __forceinline__ double blk_dist2(const BlkAcc &acc, int n, const z2 *X, double *red) {
  const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
  double s = 0;
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      int i = ty + 16 * r, j = tx + 16 * c;
      if (i < n && j < n) {
        z2 d = zsub(acc.v[r][c], X[i * n + j]);
        s += d.x * d.x + d.y * d.y;
      }
    }
  __syncthreads();
  red[tid] = s;
  __syncthreads();
  for (int w = BO_T / 2; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  return red[0];
}

}  // namespace ned
