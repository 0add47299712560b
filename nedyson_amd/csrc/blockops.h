// Workgroup-level operations on nao x nao complex FP64 blocks (nao <= 64):
// the building block of the Dyson-step history sums.
//
// The output block lives in registers in the layout of the FP64 matrix core
// (v_mfma_f64_16x16x4f64): wave w owns rows 16w..16w+15, and lane l of it
// holds, for each 16-column tile c and each q = 0..3, the element
//   row 16w + 4q + l/16, column 16c + l%16
// (real and imaginary parts in separate 4-vectors).  Products run on the
// matrix core with operand K-slices of 16 staged through LDS; tiles that lie
// wholly outside nao x nao are skipped.
#pragma once
#include "common.h"

namespace ned {

constexpr int BO_T = 256;   // threads per workgroup
constexpr int BO_KT = 16;   // K slice
constexpr int BO_NMAX = 64; // largest supported nao

typedef double blk_v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int blk_i(int q) { return 16 * (threadIdx.x >> 6) + 4 * q + ((threadIdx.x & 63) >> 4); }
__device__ __forceinline__ int blk_j(int c) { return 16 * c + (threadIdx.x & 15); }

__device__ __forceinline__ blk_v4d mfma_f64(double a, double b, blk_v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

struct BlkAcc {
  blk_v4d re[4], im[4];  // [column tile][q]
  __device__ void zero() {
#pragma unroll
    for (int c = 0; c < 4; c++) re[c] = im[c] = blk_v4d{0.0, 0.0, 0.0, 0.0};
  }
  __device__ z2 get(int q, int c) const { return z2{re[c][q], im[c][q]}; }
  __device__ void set(int q, int c, z2 v) {
    re[c][q] = v.x;
    im[c][q] = v.y;
  }
  __device__ void add(int q, int c, z2 v) {
    re[c][q] += v.x;
    im[c][q] += v.y;
  }
  // column tile c += a b over one 4-deep step (lane's fragment values a, b):
  // four real products
  __device__ __forceinline__ void mma(const z2 &a, const z2 &b, int c) {
    re[c] = mfma_f64(a.x, b.x, re[c]);
    re[c] = mfma_f64(-a.y, b.y, re[c]);
    im[c] = mfma_f64(a.x, b.y, im[c]);
    im[c] = mfma_f64(a.y, b.x, im[c]);
  }
};

// The same block accumulated with three real products per complex one
// (Gauss's 3M form): t1 = sum Re a Re b, t2 = sum Im a Im b,
// t3 = sum (Re a + Im a)(Re b + Im b), so that Re = t1 - t2 and
// Im = t3 - t1 - t2.  A quarter fewer matrix-core passes than BlkAcc, for the
// long history sums, which the FP64 matrix core bounds; the imaginary part's
// rounding error grows to about eps (|Re a| + |Im a|)(|Re b| + |Im b|) per
// term, twice the four-product form's at most.
struct BlkAcc3 {
  blk_v4d t1[4], t2[4], t3[4];
  __device__ void zero() {
#pragma unroll
    for (int c = 0; c < 4; c++) t1[c] = t2[c] = t3[c] = blk_v4d{0.0, 0.0, 0.0, 0.0};
  }
  // start from the value v
  __device__ void init(const BlkAcc &v) {
#pragma unroll
    for (int c = 0; c < 4; c++) {
      t1[c] = v.re[c];
      t2[c] = blk_v4d{0.0, 0.0, 0.0, 0.0};
      t3[c] = v.re[c] + v.im[c];
    }
  }
  __device__ BlkAcc value() const {
    BlkAcc v;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      v.re[c] = t1[c] - t2[c];
      v.im[c] = t3[c] - t1[c] - t2[c];
    }
    return v;
  }
  // element (q, c) += z
  __device__ void add(int q, int c, z2 z) {
    t1[c][q] += z.x;
    t3[c][q] += z.x + z.y;
  }
  __device__ __forceinline__ void mma(const z2 &a, const z2 &b, int c) {
    t1[c] = mfma_f64(a.x, b.x, t1[c]);
    t2[c] = mfma_f64(a.y, b.y, t2[c]);
    t3[c] = mfma_f64(a.x + a.y, b.x + b.y, t3[c]);
  }
};

// Both operands' K-slices are held with K innermost (X[row][k], Y[col][k])
// in rows of KT + 1: an MFMA fragment read (16 rows or columns, one k each
// per lane group) then reads 16 rows at the same column, which the LDS
// serves without conflicts; a k-major Y (16 consecutive columns of one k per
// lane group) measured four times slower.  KT = 16 for blk_mac (one stage);
// blk_mac2 takes two stages of 8 (BlkLds2<8>).
template <int KT>
struct BlkLdsT {
  static constexpr int kt = KT;
  z2 X[BO_NMAX][KT + 1];
  z2 Y[BO_NMAX][KT + 1];
};
using BlkLds = BlkLdsT<BO_KT>;

// op: 0 = as stored, 1 = adjoint (conjugate transpose), 2 = transpose;
// ld = row stride of the stored matrix (complex FP64 or FP32 elements)
template <typename E>
__device__ __forceinline__ z2 blk_get(const E *M, int ld, int op, int r, int c) {
  if (op == 0) return ldz(M + r * ld + c);
  z2 v = ldz(M + c * ld + r);
  return op == 1 ? zconj(v) : v;
}

// The same element left in its storage type, for operands loaded ahead of
// their use: widening (or conjugating) a value right after its load would
// make the wave wait for that load there, before the work the load is meant
// to overlap.  widen() and blk_op() finish it where it is consumed.
template <typename E>
__device__ __forceinline__ E blk_raw(const E *M, int ld, int op, int r, int c) {
  return op == 0 ? M[r * ld + c] : M[c * ld + r];
}
__device__ __forceinline__ z2 widen(z2 v) { return v; }
__device__ __forceinline__ z2 widen(float2 v) { return z2{(double)v.x, (double)v.y}; }
__device__ __forceinline__ z2 blk_op(z2 v, int op) { return op == 1 ? zconj(v) : v; }
__device__ __forceinline__ z2 zero_of(const z2 *) { return z2{0.0, 0.0}; }
__device__ __forceinline__ float2 zero_of(const float2 *) { return float2{0.0f, 0.0f}; }

// acc += coef * op(X) * op(Y)   (all threads of the workgroup call this);
// ldX, ldY are the row strides of X and Y as stored (n for a dense block).
// The next K-slice of both operands is loaded into registers while the
// current one is multiplied out of LDS (one stage, two barriers a slice: the
// form of the one-off products; the long sums use blk_mac2).
// Acc: BlkAcc (four real products per complex one) or BlkAcc3 (three).
// Lds: BlkLdsT<KT>, K slices of KT.
template <typename Acc, typename Lds, typename EX, typename EY>
__device__ __forceinline__ void blk_mac(Acc &acc, Lds &s, int n, const EX *X, int opX, const EY *Y, int opY,
                                        z2 coef, int ldX = 0, int ldY = 0) {
  constexpr int KT = Lds::kt;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (ldX == 0) ldX = n;
  if (ldY == 0) ldY = n;
  const bool wave_on = 16 * w < n;
  constexpr int PER = BO_NMAX * KT / BO_T;  // staged elements per thread and operand
  EX px[PER];  // the next K slice, in the operands' storage types
  EY py[PER];
  // element e of a slice -> (row, k) of X and (k, col) of Y, chosen so that
  // consecutive threads read consecutive addresses of the stored matrix
  // (along k when it is used as stored, along the other index when it is
  // used transposed or adjoint)
  auto xmap = [&](int e, int &r, int &kk) __attribute__((always_inline)) {
    if (opX == 0) { r = e / KT; kk = e % KT; } else { kk = e / BO_NMAX; r = e % BO_NMAX; }
  };
  auto ymap = [&](int e, int &kk, int &c) __attribute__((always_inline)) {
    if (opY == 0) { kk = e / BO_NMAX; c = e % BO_NMAX; } else { c = e / KT; kk = e % KT; }
  };
  auto getx = [&](int e, int k0) __attribute__((always_inline)) {
    int r, kk;
    xmap(e, r, kk);
    return (r < n && k0 + kk < n) ? blk_raw(X, ldX, opX, r, k0 + kk) : zero_of(X);
  };
  auto gety = [&](int e, int k0) __attribute__((always_inline)) {
    int kk, c;
    ymap(e, kk, c);
    return (c < n && k0 + kk < n) ? blk_raw(Y, ldY, opY, k0 + kk, c) : zero_of(Y);
  };
  auto fetch = [&](int k0) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int e = tid + BO_T * i;
      px[i] = getx(e, k0);
      py[i] = gety(e, k0);
    }
  };
  __syncthreads();  // operands written by this workgroup just before the call are visible
  fetch(0);
  for (int k0 = 0; k0 < n; k0 += KT) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int e = tid + BO_T * i;
      int r, kk, kk2, c;
      xmap(e, r, kk);
      ymap(e, kk2, c);
      s.X[r][kk] = zmul(coef, blk_op(widen(px[i]), opX));
      s.Y[c][kk2] = blk_op(widen(py[i]), opY);
    }
    __syncthreads();
    if (k0 + KT < n) fetch(k0 + KT);
    if (wave_on) {
      const int ksteps = min(KT, n - k0);
      for (int ks = 0; ks < ksteps; ks += 4) {
        const int kk = ks + (lane >> 4);
        const z2 a = s.X[16 * w + (lane & 15)][kk];
#pragma unroll
        for (int c = 0; c < 4; c++) {
          if (16 * c < n) {
            const z2 b = s.Y[16 * c + (lane & 15)][kk];
            acc.mma(a, b, c);
          }
        }
      }
    }
  }
}

// Two LDS stages of KT-deep slices for blk_mac2.
template <int KT>
struct BlkLds2 {
  static constexpr int kt = KT;
  z2 X[2][BO_NMAX][KT + 1];
  z2 Y[2][BO_NMAX][KT + 1];
};

// blk_mac in two LDS stages: the next slice of both operands is fetched into
// registers while the current one is multiplied out of its stage, then
// written into the idle stage; one barrier per slice instead of two, and a
// whole slice's products between a load and its use.
template <int NF = 0, typename Acc, typename Lds, typename EX, typename EY>
__device__ __forceinline__ void blk_mac2(Acc &acc, Lds &s, int n_, const EX *X, int opX, const EY *Y, int opY,
                                         z2 coef) {
  constexpr int KT = Lds::kt;
  const int n = NF ? NF : n_;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const bool wave_on = 16 * w < n;
  constexpr int PER = BO_NMAX * KT / BO_T;
  EX px[PER];
  EY py[PER];
  auto xmap = [&](int e, int &r, int &kk) __attribute__((always_inline)) {
    if (opX == 0) { r = e / KT; kk = e % KT; } else { kk = e / BO_NMAX; r = e % BO_NMAX; }
  };
  auto ymap = [&](int e, int &kk, int &c) __attribute__((always_inline)) {
    if (opY == 0) { kk = e / BO_NMAX; c = e % BO_NMAX; } else { c = e / KT; kk = e % KT; }
  };
  auto fetch = [&](int k0) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int e = tid + BO_T * i;
      int r, kk, kk2, c;
      xmap(e, r, kk);
      ymap(e, kk2, c);
      px[i] = (r < n && k0 + kk < n) ? blk_raw(X, n, opX, r, k0 + kk) : zero_of(X);
      py[i] = (c < n && k0 + kk2 < n) ? blk_raw(Y, n, opY, k0 + kk2, c) : zero_of(Y);
    }
  };
  auto stage = [&](int st) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int e = tid + BO_T * i;
      int r, kk, kk2, c;
      xmap(e, r, kk);
      ymap(e, kk2, c);
      s.X[st][r][kk] = zmul(coef, blk_op(widen(px[i]), opX));
      s.Y[st][c][kk2] = blk_op(widen(py[i]), opY);
    }
  };
  __syncthreads();  // operands written just before the call are visible; the last call's stages are free
  fetch(0);
  stage(0);
  __syncthreads();
  int st = 0;
  for (int k0 = 0; k0 < n; k0 += KT) {
    const bool more = k0 + KT < n;
    if (more) fetch(k0 + KT);
    if (wave_on) {
      const int ksteps = min(KT, n - k0);
      for (int ks = 0; ks < ksteps; ks += 4) {
        const int kk = ks + (lane >> 4);
        const z2 a = s.X[st][16 * w + (lane & 15)][kk];
#pragma unroll
        for (int c = 0; c < 4; c++) {
          if (16 * c < n) {
            const z2 b = s.Y[st][16 * c + (lane & 15)][kk];
            acc.mma(a, b, c);
          }
        }
      }
    }
    if (more) stage(st ^ 1);
    __syncthreads();
    st ^= 1;
  }
}

// acc += coef * op(X)   (elementwise, no product; Acc: BlkAcc or BlkAcc3)
template <typename Acc, typename E>
__device__ __forceinline__ void blk_axpy(Acc &acc, int n, const E *X, int opX, z2 coef) {
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      int i = blk_i(q), j = blk_j(c);
      if (i < n && j < n) acc.add(q, c, zmul(coef, blk_get(X, n, opX, i, j)));
    }
}

__device__ __forceinline__ void blk_load(BlkAcc &acc, int n, const z2 *X) {
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      int i = blk_i(q), j = blk_j(c);
      acc.set(q, c, (i < n && j < n) ? X[i * n + j] : z2{0.0, 0.0});
    }
}

__device__ __forceinline__ void blk_store(const BlkAcc &acc, int n, z2 *X) {
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      int i = blk_i(q), j = blk_j(c);
      if (i < n && j < n) X[i * n + j] = acc.get(q, c);
    }
}

__device__ __forceinline__ void blk_scale(BlkAcc &acc, z2 f) {
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc.set(q, c, zmul(acc.get(q, c), f));
}

// acc = diag(d) (the identity times d)
__device__ __forceinline__ void blk_diag(BlkAcc &acc, int n, z2 d) {
  acc.zero();
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int c = 0; c < 4; c++)
      if (blk_i(q) == blk_j(c) && blk_i(q) < n) acc.set(q, c, d);
}

// squared Frobenius distance between acc and X, reduced over the workgroup
// (result valid in thread 0)
__device__ __forceinline__ double blk_dist2(const BlkAcc &acc, int n, const z2 *X, double *red) {
  const int tid = threadIdx.x;
  double s = 0;
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      int i = blk_i(q), j = blk_j(c);
      if (i < n && j < n) {
        z2 d = zsub(acc.get(q, c), X[i * n + j]);
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

namespace ned {

// ---------------------------------------------------------------- tiles
// One 16 x 16 output tile per wave at (r0, c0) (r0, c0 multiples of 16), in
// the same lane layout as BlkAcc.  Used where one block is split over several
// workgroups: the retarded recurrence is separable by rows and the lesser one
// by columns, so a 16-row (16-column) band of a block runs on its own
// workgroup with its four waves on the four column (row) tiles.
struct TileAcc {
  blk_v4d re, im;
  __device__ void zero() { re = im = blk_v4d{0.0, 0.0, 0.0, 0.0}; }
  __device__ z2 get(int q) const { return z2{re[q], im[q]}; }
  __device__ void set(int q, z2 v) {
    re[q] = v.x;
    im[q] = v.y;
  }
  __device__ __forceinline__ void mma(const z2 &x, const z2 &y) {
    re = mfma_f64(x.x, y.x, re);
    im = mfma_f64(x.x, y.y, im);
    re = mfma_f64(-x.y, y.y, re);
    im = mfma_f64(x.y, y.x, im);
  }
};

// one tile in the three-product form of BlkAcc3
struct TileAcc3 {
  blk_v4d t1, t2, t3;
  __device__ void zero() { t1 = t2 = t3 = blk_v4d{0.0, 0.0, 0.0, 0.0}; }
  __device__ void init(const TileAcc &v) {
    t1 = v.re;
    t2 = blk_v4d{0.0, 0.0, 0.0, 0.0};
    t3 = v.re + v.im;
  }
  __device__ TileAcc value() const {
    TileAcc v;
    v.re = t1 - t2;
    v.im = t3 - t1 - t2;
    return v;
  }
  // element q += z
  __device__ void add(int q, z2 z) {
    t1[q] += z.x;
    t3[q] += z.x + z.y;
  }
  __device__ __forceinline__ void mma(const z2 &x, const z2 &y) {
    t1 = mfma_f64(x.x, y.x, t1);
    t2 = mfma_f64(x.y, y.y, t2);
    t3 = mfma_f64(x.x + x.y, y.x + y.y, t3);
  }
};

__device__ __forceinline__ int tile_i(int r0, int q) { return r0 + 4 * q + ((threadIdx.x & 63) >> 4); }
__device__ __forceinline__ int tile_j(int c0) { return c0 + (threadIdx.x & 15); }

__device__ __forceinline__ void tile_load(TileAcc &acc, int n, const z2 *X, int r0, int c0) {
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int i = tile_i(r0, q), j = tile_j(c0);
    acc.set(q, (i < n && j < n) ? X[i * n + j] : z2{0.0, 0.0});
  }
}

__device__ __forceinline__ void tile_store(const TileAcc &acc, int n, z2 *X, int r0, int c0) {
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int i = tile_i(r0, q), j = tile_j(c0);
    if (i < n && j < n) X[i * n + j] = acc.get(q);
  }
}

// Operand fragments of v_mfma_f64_16x16x4f64 for one wave's 16 x 16 tile,
// read straight from memory: no LDS staging and no barrier, so a product
// whose operands are read once costs one memory round trip plus the matrix
// core time.  This is the form for the dependency chains, where one
// workgroup's latency is the whole cost.  Step s covers k = 4s..4s+3; lane l
// holds
//   left  operand: op(L)[r0 + l%16][4s + l/16]
//   right operand: op(R)[4s + l/16][c0 + l%16]
// and zero outside nao x nao.
// A fragment may hold only the NS steps from s0 on (half the K range in
// throughput kernels, which keeps registers for occupancy).
// E: the element type the fragment is loaded from; an FP32 fragment is
// widened where frag_mac uses it, so its loads overlap whatever comes between
// (FP32 fragments are read as stored, op 0)
template <int NS = BO_NMAX / 4, typename E = z2>
struct FragT {
  E v[NS];
};
using Frag = FragT<>;

template <int NS, typename E>
__device__ __forceinline__ void frag_left(FragT<NS, E> &f, int n, const E *L, int opL, int r0, int s0 = 0) {
  const int lane = threadIdx.x & 63, r = r0 + (lane & 15);
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int kk = 4 * (s0 + s) + (lane >> 4);
    if constexpr (sizeof(E) == sizeof(z2)) f.v[s] = (r < n && kk < n) ? blk_get(L, n, opL, r, kk) : z2{0.0, 0.0};
    else f.v[s] = (r < n && kk < n) ? blk_raw(L, n, 0, r, kk) : zero_of(L);
  }
}

template <int NS, typename E>
__device__ __forceinline__ void frag_right(FragT<NS, E> &f, int n, const E *R, int opR, int c0, int s0 = 0) {
  const int lane = threadIdx.x & 63, c = c0 + (lane & 15);
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int kk = 4 * (s0 + s) + (lane >> 4);
    if constexpr (sizeof(E) == sizeof(z2)) f.v[s] = (kk < n && c < n) ? blk_get(R, n, opR, kk, c) : z2{0.0, 0.0};
    else f.v[s] = (kk < n && c < n) ? blk_raw(R, n, 0, kk, c) : zero_of(R);
  }
}

// acc += coef * L R over the fragments' steps; coef scales the left operand,
// as the LDS-staged products do, so both forms round alike
// (Acc: TileAcc or TileAcc3)
template <int NS, typename Acc, typename EA, typename EB>
__device__ __forceinline__ void frag_mac(Acc &acc, const FragT<NS, EA> &a, const FragT<NS, EB> &b, int n, z2 coef,
                                         int s0 = 0) {
#pragma unroll
  for (int s = 0; s < NS; s++) {
    if (4 * (s0 + s) < n) acc.mma(zmul(coef, widen(a.v[s])), widen(b.v[s]));
  }
}

// the four tiles of a 16-row band (tile (0, 16w) from wave w) as the left
// operand of the next product: band[16][NMAX + 1] in LDS
__device__ __forceinline__ void frag_left_band(Frag &f, const TileAcc &acc, z2 (*band)[BO_NMAX + 1], int n) {
  const int lane = threadIdx.x & 63, c0 = 16 * (threadIdx.x >> 6);
#pragma unroll
  for (int q = 0; q < 4; q++) band[4 * q + (lane >> 4)][c0 + (lane & 15)] = acc.get(q);
  __syncthreads();
#pragma unroll
  for (int s = 0; s < BO_NMAX / 4; s++) {
    const int kk = 4 * s + (lane >> 4);
    f.v[s] = kk < n ? band[lane & 15][kk] : z2{0.0, 0.0};
  }
}

// the four tiles of a 16-column band (tile (16w, 0) from wave w) as the
// right operand of the next product: band[NMAX][16 + 1] in LDS
__device__ __forceinline__ void frag_right_band(Frag &f, const TileAcc &acc, z2 (*band)[17], int n) {
  const int lane = threadIdx.x & 63, r0 = 16 * (threadIdx.x >> 6);
#pragma unroll
  for (int q = 0; q < 4; q++) band[r0 + 4 * q + (lane >> 4)][lane & 15] = acc.get(q);
  __syncthreads();
#pragma unroll
  for (int s = 0; s < BO_NMAX / 4; s++) {
    const int kk = 4 * s + (lane >> 4);
    f.v[s] = kk < n ? band[kk][lane & 15] : z2{0.0, 0.0};
  }
}

// sum of v over the workgroup, valid in thread 0 (red: 4 doubles of LDS)
__device__ __forceinline__ double wg_sum(double v, double *red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}


}  // namespace ned
