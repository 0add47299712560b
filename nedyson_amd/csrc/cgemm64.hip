// Batched 64 x 64 complex products of the self-energy chain in split-bf16
// (three v_mfma_f32_16x16x32_bf16 per real product, FP32 accumulation), for
// nao = 64.  One workgroup takes one time point p and a run of G auxiliary
// indices; the operand that all G products share is split and staged in LDS
// once, the per-index operand is streamed, and results leave either as FP32
// or already split into the bf16 planes that the next GEMM reads, in the
// layout it wants (transposed through LDS where needed).  These products move
// ~16 bytes per 64 flops, so they are HBM-bound; the point of the kernel is to
// touch every byte once.
//
// Operands are K-contiguous: A as [row][k], B as [col][k], so that a lane's
// eight consecutive K values are one 16-byte LDS read.
//
//   W^a = V^a A_p     -> W (FP32, [p][pl][a][i][k]) and Wt (split, [p][pl][a][k][i])
//   Z^b = B_p^T V^b   -> Zt (split, [p][pl][b][k][i])          (V^b symmetric)
//   C^a = B_p W^a     -> rows (p, pl, a) of the exchange operand (split, [i][k])
//   E_p = sum_a W^a H^a over the run of a -> FP32 partial sums [p][s][pl][i][j]
#include "common.h"
#include "cgemm64.h"

namespace ned {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

constexpr int N64 = 64, KS = 72;  // padded K stride (144 bytes: 16-byte aligned, conflict-free columns)

// LDS image of one 64 x 64 operand: [plane][row][k], hi and lo bf16 parts
struct Img {
  uint16_t h[2][N64][KS];
  uint16_t l[2][N64][KS];
};

__device__ __forceinline__ f32x4 mma3(u32x4_t ah, u32x4_t al, u32x4_t bh, u32x4_t bl, f32x4 c) {
  const bf16x8 a_h = __builtin_bit_cast(bf16x8, ah), a_l = __builtin_bit_cast(bf16x8, al);
  const bf16x8 b_h = __builtin_bit_cast(bf16x8, bh), b_l = __builtin_bit_cast(bf16x8, bl);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_l, b_h, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h, b_l, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h, b_h, c, 0, 0, 0);
  return c;
}

__device__ __forceinline__ u32x4_t negbf(u32x4_t v) { return v ^ 0x80008000u; }

// thread tid's share of an FP32 [row][k] plane: four float4, element
// e = tid + 256 it at row e/16, columns 4(e%16)..+3
__device__ __forceinline__ int f32_row(int it) { return (threadIdx.x + 256 * it) >> 4; }
__device__ __forceinline__ int f32_col(int it) { return ((threadIdx.x + 256 * it) & 15) * 4; }

// split four FP32 values into the image's hi and lo planes
__device__ __forceinline__ void put_f32(Img &im, int pl, int it, float4 v) {
  const int row = f32_row(it), c4 = f32_col(it);
  const float f[4] = {v.x, v.y, v.z, v.w};
  uint16_t hh[4], ll[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    hh[q] = f2bf(f[q]);
    ll[q] = f2bf(f[q] - bf2f(hh[q]));
  }
  *reinterpret_cast<uint2 *>(&im.h[pl][row][c4]) =
      uint2{(uint32_t)hh[0] | ((uint32_t)hh[1] << 16), (uint32_t)hh[2] | ((uint32_t)hh[3] << 16)};
  *reinterpret_cast<uint2 *>(&im.l[pl][row][c4]) =
      uint2{(uint32_t)ll[0] | ((uint32_t)ll[1] << 16), (uint32_t)ll[2] | ((uint32_t)ll[3] << 16)};
}

// FP32 [row][k] planes (plane stride ps) -> split image
__device__ __forceinline__ void load_f32(Img &im, const float *src, long long ps, int planes) {
  for (int pl = 0; pl < planes; pl++)
#pragma unroll
    for (int it = 0; it < 4; it++)
      put_f32(im, pl, it, *reinterpret_cast<const float4 *>(src + pl * ps + f32_row(it) * N64 + f32_col(it)));
}

// pre-split [row][k] planes -> image
__device__ __forceinline__ void load_bf(Img &im, const uint16_t *hi, const uint16_t *lo, long long ps, int planes) {
  const int tid = threadIdx.x;
  for (int pl = 0; pl < planes; pl++)
#pragma unroll
    for (int it = 0; it < 2; it++) {
      const int e = tid + 256 * it, row = e >> 3, c8 = (e & 7) * 8;
      *reinterpret_cast<uint4 *>(&im.h[pl][row][c8]) = *reinterpret_cast<const uint4 *>(hi + pl * ps + row * N64 + c8);
      *reinterpret_cast<uint4 *>(&im.l[pl][row][c8]) = *reinterpret_cast<const uint4 *>(lo + pl * ps + row * N64 + c8);
    }
}

// acc[j][pl] += A(rows 16w..16w+15) * B(cols 16j..16j+15), complex or real
// operands by plane count
template <int PA, int PB>
__device__ __forceinline__ void mac64(f32x4 (&acc)[4][2], const Img &A, const Img &B) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int ar = 16 * w + (lane & 15);
#pragma unroll
  for (int ks = 0; ks < N64; ks += 32) {
    const int kq = ks + 8 * (lane >> 4);
    u32x4_t ah[PA], al[PA];
#pragma unroll
    for (int pl = 0; pl < PA; pl++) {
      ah[pl] = *reinterpret_cast<const u32x4_t *>(&A.h[pl][ar][kq]);
      al[pl] = *reinterpret_cast<const u32x4_t *>(&A.l[pl][ar][kq]);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int bc = 16 * j + (lane & 15);
      u32x4_t bh[PB], bl[PB];
#pragma unroll
      for (int pl = 0; pl < PB; pl++) {
        bh[pl] = *reinterpret_cast<const u32x4_t *>(&B.h[pl][bc][kq]);
        bl[pl] = *reinterpret_cast<const u32x4_t *>(&B.l[pl][bc][kq]);
      }
      if constexpr (PA == 2 && PB == 2) {
        acc[j][0] = mma3(ah[0], al[0], bh[0], bl[0], acc[j][0]);
        acc[j][0] = mma3(negbf(ah[1]), negbf(al[1]), bh[1], bl[1], acc[j][0]);
        acc[j][1] = mma3(ah[0], al[0], bh[1], bl[1], acc[j][1]);
        acc[j][1] = mma3(ah[1], al[1], bh[0], bl[0], acc[j][1]);
      } else if constexpr (PA == 1) {
        acc[j][0] = mma3(ah[0], al[0], bh[0], bl[0], acc[j][0]);
        acc[j][1] = mma3(ah[0], al[0], bh[1], bl[1], acc[j][1]);
      } else {
        acc[j][0] = mma3(ah[0], al[0], bh[0], bl[0], acc[j][0]);
        acc[j][1] = mma3(ah[1], al[1], bh[0], bl[0], acc[j][1]);
      }
    }
  }
}

__device__ __forceinline__ void zero(f32x4 (&acc)[4][2]) {
#pragma unroll
  for (int j = 0; j < 4; j++) acc[j][0] = acc[j][1] = f32x4{0, 0, 0, 0};
}

// split the accumulators into an image, natural ([i][k]) or transposed ([k][i])
__device__ __forceinline__ void stage(Img &st, const f32x4 (&acc)[4][2], bool transposed) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (transposed) {
    // a lane's four values q are consecutive i of one row k of the image:
    // one 8-byte write per plane and part instead of four 2-byte ones
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int pl = 0; pl < 2; pl++) {
        const int i0 = 16 * w + 4 * (lane >> 4), k = 16 * j + (lane & 15);
        uint16_t h[4], l[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const float v = acc[j][pl][q];
          h[q] = f2bf(v);
          l[q] = f2bf(v - bf2f(h[q]));
        }
        *reinterpret_cast<uint2 *>(&st.h[pl][k][i0]) =
            uint2{(uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16)};
        *reinterpret_cast<uint2 *>(&st.l[pl][k][i0]) =
            uint2{(uint32_t)l[0] | ((uint32_t)l[1] << 16), (uint32_t)l[2] | ((uint32_t)l[3] << 16)};
      }
    return;
  }
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int pl = 0; pl < 2; pl++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int i = 16 * w + 4 * (lane >> 4) + q, k = 16 * j + (lane & 15);
        const float v = acc[j][pl][q];
        const uint16_t h = f2bf(v), l = f2bf(v - bf2f(h));
        if (transposed) { st.h[pl][k][i] = h; st.l[pl][k][i] = l; }
        else { st.h[pl][i][k] = h; st.l[pl][i][k] = l; }
      }
}

// image rows -> global [row][k] planes with row stride ld and plane stride ps
__device__ __forceinline__ void store_bf(const Img &st, uint16_t *hi, uint16_t *lo, long long ld, long long ps) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int pl = 0; pl < 2; pl++)
#pragma unroll
    for (int it = 0; it < 2; it++) {
      const int e = tid + 256 * it, row = e >> 3, c8 = (e & 7) * 8;
      *reinterpret_cast<uint4 *>(hi + pl * ps + row * ld + c8) = *reinterpret_cast<const uint4 *>(&st.h[pl][row][c8]);
      *reinterpret_cast<uint4 *>(lo + pl * ps + row * ld + c8) = *reinterpret_cast<const uint4 *>(&st.l[pl][row][c8]);
    }
}

__device__ __forceinline__ void store_f32(float *dst, long long ps, const f32x4 (&acc)[4][2]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int pl = 0; pl < 2; pl++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int i = 16 * w + 4 * (lane >> 4) + q, k = 16 * j + (lane & 15);
        dst[pl * ps + i * N64 + k] = acc[j][pl][q];
      }
}

constexpr long long E2 = (long long)N64 * N64;  // elements of one plane

// W^a = V^a A_p: A = V^a (real, pre-split [a][i][l]), B = A_p^T ([k][l], FP32)
__global__ void __launch_bounds__(256) kern_c64_w(const uint16_t *__restrict__ Vh, const uint16_t *__restrict__ Vl,
                                                  const float *__restrict__ At, int a, int G, float *__restrict__ W,
                                                  uint16_t *__restrict__ Wth, uint16_t *__restrict__ Wtl) {
  __shared__ __attribute__((aligned(16))) Img S, X;
  const int p = blockIdx.y, a0 = blockIdx.x * G;
  const long long an2 = (long long)a * E2;
  load_f32(S, At + (size_t)p * 2 * E2, E2, 2);
  for (int x = a0; x < a0 + G && x < a; x++) {
    __syncthreads();
    load_bf(X, Vh + (size_t)x * E2, Vl + (size_t)x * E2, 0, 1);
    __syncthreads();
    f32x4 acc[4][2];
    zero(acc);
    mac64<1, 2>(acc, X, S);
    const size_t o = (size_t)p * 2 * an2 + (size_t)x * E2;
    if (W) store_f32(W + o, an2, acc);  // only the split form when E reads that
    __syncthreads();
    stage(X, acc, true);
    __syncthreads();
    store_bf(X, Wth + o, Wtl + o, N64, an2);
  }
}

// Z^b = B_p^T V^b: A = B_p^T ([i][l], FP32), B = V^b ([k][l] = V^b[l][k], pre-split)
__global__ void __launch_bounds__(256) kern_c64_z(const uint16_t *__restrict__ Vh, const uint16_t *__restrict__ Vl,
                                                  const float *__restrict__ Bt, int a, int G,
                                                  uint16_t *__restrict__ Zth, uint16_t *__restrict__ Ztl) {
  __shared__ __attribute__((aligned(16))) Img S, X;
  const int p = blockIdx.y, a0 = blockIdx.x * G;
  const long long an2 = (long long)a * E2;
  load_f32(S, Bt + (size_t)p * 2 * E2, E2, 2);
  for (int x = a0; x < a0 + G && x < a; x++) {
    __syncthreads();
    load_bf(X, Vh + (size_t)x * E2, Vl + (size_t)x * E2, 0, 1);
    __syncthreads();
    f32x4 acc[4][2];
    zero(acc);
    mac64<2, 1>(acc, S, X);
    __syncthreads();
    stage(X, acc, true);
    __syncthreads();
    const size_t o = (size_t)p * 2 * an2 + (size_t)x * E2;
    store_bf(X, Zth + o, Ztl + o, N64, an2);
  }
}

// C^a = B_p W^a: A = B_p ([i][l], FP32), B = Wt^a ([k][l], pre-split); rows
// (p, pl, a) of the exchange operand, row stride ld
// With SYM the row is written in the split layout of the exchange: groups of
// 8 pairs from the table (m << 8 | q), pairs m <= q as X_mq + X_qm (X_mm on the
// diagonal) in columns [0, ns), pairs m < q as X_mq - X_qm from column ks, and
// zeros in the 32 padding columns after the Pi block and after the pairs m < q.
template <bool SYM>
__global__ void __launch_bounds__(256) kern_c64_c(const float *__restrict__ Bm, const uint16_t *__restrict__ Wth,
                                                  const uint16_t *__restrict__ Wtl, int a, int G,
                                                  uint16_t *__restrict__ Sh, uint16_t *__restrict__ Sl, long long ld,
                                                  const uint16_t *__restrict__ pairs, long long ks) {
  __shared__ __attribute__((aligned(16))) Img S, X;
  const int p = blockIdx.y, a0 = blockIdx.x * G;
  const long long an2 = (long long)a * E2;
  load_f32(S, Bm + (size_t)p * 2 * E2, E2, 2);
  for (int x = a0; x < a0 + G && x < a; x++) {
    __syncthreads();
    const size_t o = (size_t)p * 2 * an2 + (size_t)x * E2;
    load_bf(X, Wth + o, Wtl + o, an2, 2);
    __syncthreads();
    f32x4 acc[4][2];
    zero(acc);
    mac64<2, 2>(acc, S, X);
    __syncthreads();
    const size_t r0 = ((size_t)p * 2 * a + x) * ld;  // plane 0 row; plane 1 is a rows further
    const int tid = threadIdx.x;
    if constexpr (SYM) {
      constexpr int NS = N64 * (N64 + 1) / 2, NA = N64 * (N64 - 1) / 2, NS8 = NS / 8;
      // fp32 image of the block (rows m, columns q, stride 65 words) in X's space
      float(*F)[N64][65] = reinterpret_cast<float(*)[N64][65]>(&X);
      {
        const int lane = tid & 63, w = tid >> 6;
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
          for (int pl = 0; pl < 2; pl++)
#pragma unroll
            for (int q = 0; q < 4; q++) F[pl][16 * w + 4 * (lane >> 4) + q][16 * j + (lane & 15)] = acc[j][pl][q];
      }
      __syncthreads();
#pragma unroll
      for (int it = 0; it < 2; it++) {
        const int g = tid + 256 * it;  // 512 groups: NS8 symmetric, then NA / 8 antisymmetric
        const uint4 pk = *reinterpret_cast<const uint4 *>(pairs + 8 * g);
        const uint32_t pw[4] = {pk.x, pk.y, pk.z, pk.w};
        const size_t col = g < NS8 ? (size_t)8 * g : (size_t)ks + 8 * (g - NS8);
#pragma unroll
        for (int pl = 0; pl < 2; pl++) {
          uint32_t hw[4], lw[4];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            hw[u] = 0;
            lw[u] = 0;
#pragma unroll
            for (int hf = 0; hf < 2; hf++) {
              const uint32_t e = (pw[u] >> (16 * hf)) & 0xffffu;
              const int m = (int)(e >> 8), q = (int)(e & 255u);
              const float xv = F[pl][m][q], yv = F[pl][q][m];
              const float v = g < NS8 ? (m == q ? xv : xv + yv) : xv - yv;
              const uint16_t h = f2bf(v), l = f2bf(v - bf2f(h));
              hw[u] |= (uint32_t)h << (16 * hf);
              lw[u] |= (uint32_t)l << (16 * hf);
            }
          }
          const size_t go = r0 + (size_t)pl * a * ld + col;
          *reinterpret_cast<uint4 *>(Sh + go) = make_uint4(hw[0], hw[1], hw[2], hw[3]);
          *reinterpret_cast<uint4 *>(Sl + go) = make_uint4(lw[0], lw[1], lw[2], lw[3]);
        }
      }
      if (tid < 16) {
        const int pl = tid >> 3, w = tid & 7;
        const size_t col = w < 4 ? (size_t)(NS + a + 8 * w) : (size_t)ks + NA + 8 * (w - 4);
        const size_t go = r0 + (size_t)pl * a * ld + col;
        *reinterpret_cast<uint4 *>(Sh + go) = make_uint4(0, 0, 0, 0);
        *reinterpret_cast<uint4 *>(Sl + go) = make_uint4(0, 0, 0, 0);
      }
      continue;
    }
    stage(X, acc, false);
    __syncthreads();
    // the 64 x 64 block is one row of the operand: image row i -> columns i*64..
#pragma unroll
    for (int pl = 0; pl < 2; pl++)
#pragma unroll
      for (int it = 0; it < 2; it++) {
        const int e = tid + 256 * it, row = e >> 3, c8 = (e & 7) * 8;
        const size_t go = r0 + (size_t)pl * a * ld + row * N64 + c8;
        *reinterpret_cast<uint4 *>(Sh + go) = *reinterpret_cast<const uint4 *>(&X.h[pl][row][c8]);
        *reinterpret_cast<uint4 *>(Sl + go) = *reinterpret_cast<const uint4 *>(&X.l[pl][row][c8]);
      }
  }
}

// one FP32 value split into the image's hi and lo planes at [row][col]
__device__ __forceinline__ void put1(Img &im, int pl, int row, int col, float v) {
  const uint16_t h = f2bf(v);
  im.h[pl][row][col] = h;
  im.l[pl][row][col] = f2bf(v - bf2f(h));
}

// partial E over a in [a0, a0+G): A = W^a ([i][k], FP32), B = Ht^a ([j][k], FP32).
// With SYM, H^a comes as the two outputs of the split exchange instead: row
// (item, plane, a) of Y holds Cs over the strict pairs k < j (in the order of
// the pair table's entries [ns, ns + na)) and then the diagonal, and Ca over
// the same strict pairs from column nsp; H_kj = Cs + Ca, H_jk = Cs - Ca, H_kk = Cs.
// Thread t takes the strict pairs t + 256 u (neighbouring lanes hold
// neighbouring pairs: coalesced loads, and image writes of neighbouring rows)
// and, below 64, the diagonal entry t.
// Also with SYM, W^a comes as its transposed split (Wt[k][i] = W^a[i][k], the
// layout c64_w leaves for C and Pi): lane t of a wave takes row k = t of Wt,
// 8 columns a load, and writes them down column k of the image.
template <bool SYM>
__global__ void __launch_bounds__(256) kern_c64_e(const float *__restrict__ W, const float *__restrict__ Ht, int a,
                                                  int G, float *__restrict__ Ep, const float *__restrict__ Y,
                                                  long long ldy, int nsp, const uint16_t *__restrict__ pairs,
                                                  const uint16_t *__restrict__ Wth, const uint16_t *__restrict__ Wtl) {
  __shared__ __attribute__((aligned(16))) Img S, X;
  const int p = blockIdx.y, s = blockIdx.x, a0 = s * G, a1 = min(a0 + G, a), tid = threadIdx.x;
  const long long an2 = (long long)a * E2;
  constexpr int NS = N64 * (N64 + 1) / 2, NA = N64 * (N64 - 1) / 2, PU = (NA + 255) / 256;  // 8 pairs a thread
  uint16_t pk[PU];
  if constexpr (SYM)
#pragma unroll
    for (int u = 0; u < PU; u++) pk[u] = tid + 256 * u < NA ? pairs[NS + tid + 256 * u] : (uint16_t)0;
  // both operands are read once: the next index's are loaded into registers
  // while the current product runs
  float4 rw[2][4], rh[2][4];
  float ys[2][PU], ya[2][PU], yd[2];
  uint4 wh[2][2], wl[2][2];
  auto gload = [&](int x) __attribute__((always_inline)) {
    const size_t o = (size_t)p * 2 * an2 + (size_t)x * E2;
#pragma unroll
    for (int pl = 0; pl < 2; pl++) {
      if constexpr (!SYM)
#pragma unroll
        for (int it = 0; it < 4; it++) {
          const size_t e = o + pl * an2 + f32_row(it) * N64 + f32_col(it);
          rw[pl][it] = *reinterpret_cast<const float4 *>(W + e);
          rh[pl][it] = *reinterpret_cast<const float4 *>(Ht + e);
        }
      if constexpr (SYM) {
#pragma unroll
        for (int it = 0; it < 2; it++) {
          const int e = tid + 256 * it, k = e & 63, i0 = (e >> 6) * 8;
          const size_t g = o + pl * an2 + (size_t)k * N64 + i0;
          wh[pl][it] = *reinterpret_cast<const uint4 *>(Wth + g);
          wl[pl][it] = *reinterpret_cast<const uint4 *>(Wtl + g);
        }
        const float *y = Y + (((size_t)p * 2 + pl) * a + x) * ldy;
#pragma unroll
        for (int u = 0; u < PU; u++) {
          const int q = tid + 256 * u;
          ys[pl][u] = q < NA ? y[q] : 0.0f;
          ya[pl][u] = q < NA ? y[nsp + q] : 0.0f;
        }
        yd[pl] = tid < N64 ? y[NA + tid] : 0.0f;
      }
    }
  };
  f32x4 acc[4][2];
  zero(acc);
  if (a0 < a1) gload(a0);
  for (int x = a0; x < a1; x++) {
    __syncthreads();
#pragma unroll
    for (int pl = 0; pl < 2; pl++) {
      if constexpr (!SYM)
#pragma unroll
        for (int it = 0; it < 4; it++) {
          put_f32(S, pl, it, rw[pl][it]);
          put_f32(X, pl, it, rh[pl][it]);
        }
      if constexpr (SYM) {
#pragma unroll
        for (int it = 0; it < 2; it++) {
          const int e = tid + 256 * it, k = e & 63, i0 = (e >> 6) * 8;
          const uint32_t hw[4] = {wh[pl][it].x, wh[pl][it].y, wh[pl][it].z, wh[pl][it].w};
          const uint32_t lw[4] = {wl[pl][it].x, wl[pl][it].y, wl[pl][it].z, wl[pl][it].w};
#pragma unroll
          for (int u = 0; u < 8; u++) {
            S.h[pl][i0 + u][k] = (uint16_t)(hw[u >> 1] >> (16 * (u & 1)));
            S.l[pl][i0 + u][k] = (uint16_t)(lw[u >> 1] >> (16 * (u & 1)));
          }
        }
#pragma unroll
        for (int u = 0; u < PU; u++)
          if (tid + 256 * u < NA) {
            const int k = pk[u] >> 8, j = pk[u] & 255;
            put1(X, pl, j, k, ys[pl][u] + ya[pl][u]);  // Ht[j][k] = H_kj
            put1(X, pl, k, j, ys[pl][u] - ya[pl][u]);  // Ht[k][j] = H_jk
          }
        if (tid < N64) put1(X, pl, tid, tid, yd[pl]);
      }
    }
    __syncthreads();
    if (x + 1 < a1) gload(x + 1);
    mac64<2, 2>(acc, S, X);
  }
  store_f32(Ep + ((size_t)p * gridDim.x + s) * 2 * E2, E2, acc);
}

}  // namespace

void c64_w(const uint16_t *Vh, const uint16_t *Vl, const float *At, int P, int a, int G, float *W, uint16_t *Wth,
           uint16_t *Wtl, hipStream_t s) {
  hipLaunchKernelGGL(kern_c64_w, dim3((a + G - 1) / G, P), dim3(256), 0, s, Vh, Vl, At, a, G, W, Wth, Wtl);
  check_launch("c64_w");
}

void c64_z(const uint16_t *Vh, const uint16_t *Vl, const float *Bt, int P, int a, int G, uint16_t *Zth,
           uint16_t *Ztl, hipStream_t s) {
  hipLaunchKernelGGL(kern_c64_z, dim3((a + G - 1) / G, P), dim3(256), 0, s, Vh, Vl, Bt, a, G, Zth, Ztl);
  check_launch("c64_z");
}

void c64_c(const float *Bm, const uint16_t *Wth, const uint16_t *Wtl, int P, int a, int G, uint16_t *Sh,
           uint16_t *Sl, long long ld, hipStream_t s, const uint16_t *pairs, long long ks) {
  if (pairs)
    hipLaunchKernelGGL(kern_c64_c<true>, dim3((a + G - 1) / G, P), dim3(256), 0, s, Bm, Wth, Wtl, a, G, Sh, Sl, ld,
                       pairs, ks);
  else
    hipLaunchKernelGGL(kern_c64_c<false>, dim3((a + G - 1) / G, P), dim3(256), 0, s, Bm, Wth, Wtl, a, G, Sh, Sl, ld,
                       pairs, ks);
  check_launch("c64_c");
}

void c64_e(const float *W, const float *Ht, int P, int a, int S, float *Ep, hipStream_t s, const float *Y,
           long long ldy, int nsp, const uint16_t *pairs, const uint16_t *Wth, const uint16_t *Wtl) {
  if (a % S) throw std::invalid_argument("c64_e: naux must be a multiple of the split count");
  if (Y)
    hipLaunchKernelGGL(kern_c64_e<true>, dim3(S, P), dim3(256), 0, s, W, Ht, a, a / S, Ep, Y, ldy, nsp, pairs, Wth,
                       Wtl);
  else
    hipLaunchKernelGGL(kern_c64_e<false>, dim3(S, P), dim3(256), 0, s, W, Ht, a, a / S, Ep, Y, ldy, nsp, pairs, Wth,
                       Wtl);
  check_launch("c64_e");
}

}  // namespace ned
