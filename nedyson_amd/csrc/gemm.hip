// Strided batched GEMM on the CDNA4 matrix cores (gfx950), real or planar
// complex operands, in FP64, FP32 or split-bf16 (bf16x3) arithmetic.
//
// Work decomposition: one 256-thread workgroup (4 wave64) per 64x64 output
// tile; each wave owns a 32x32 quadrant = 2x2 MFMA 16x16 blocks.  Operand
// tiles are staged in LDS; the loader follows whichever operand dimension is
// unit-stride so that consecutive lanes read consecutive addresses.
#include <type_traits>
#include "common.h"
#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

namespace ned {

namespace {

constexpr int BM = 64, BN = 64, NT = 256;

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

__device__ __forceinline__ long long koff(const GOp &o, int k) {
  if (o.kdiv) return (long long)(k / o.kdiv) * o.ks_hi + (long long)(k % o.kdiv) * o.ks_lo;
  return (long long)k * o.ks_lo;
}
__device__ __forceinline__ long long boff(const GOp &o, int b) {
  return (long long)(b / o.bdiv) * o.bs_hi + (long long)(b % o.bdiv) * o.bs_lo;
}

// round-to-nearest-even float -> bf16 bits (finite inputs)

template <typename T>
__device__ __forceinline__ T ldg(const void *p, long long off) {
  return reinterpret_cast<const T *>(p)[off];
}

// ---------------------------------------------------------------- FP64 / FP32
template <typename T, int NPA, int NPB>
__global__ void __launch_bounds__(NT) kern_gemm_fp(GemmDesc d) {
  constexpr int BK = 16;
  constexpr int NPC = (NPA > 1 || NPB > 1) ? 2 : 1;
  __shared__ T As[NPA][BM][BK + 1];
  __shared__ T Bs[NPB][BK][BN + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int tilesN = (d.N + BN - 1) / BN;
  const int m0 = (blockIdx.x / tilesN) * BM, n0 = (blockIdx.x % tilesN) * BN;
  const bool a_kfast = d.A.ks_lo == 1 && d.A.kdiv == 0;
  const bool b_nfast = d.B.ms == 1;

  for (int b = blockIdx.y; b < d.batch; b += gridDim.y) {
    const long long ba = boff(d.A, b), bb = boff(d.B, b);
    using V = typename std::conditional<sizeof(T) == 8, v4d, v4f>::type;
    V acc[2][2][NPC];
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++)
        for (int p = 0; p < NPC; p++) acc[i][j][p] = V{0, 0, 0, 0};

    for (int k0 = 0; k0 < d.K; k0 += BK) {
      // ---- stage A (BM x BK) and B (BK x BN)
#pragma unroll
      for (int it = 0; it < (BM * BK) / NT; it++) {
        int e = tid + it * NT;
        int r, kk;
        if (a_kfast) { r = e / BK; kk = e % BK; } else { r = e % BM; kk = e / BM; }
        int gm = m0 + r, gk = k0 + kk;
        bool ok = gm < d.M && gk < d.K;
        long long off = ba + (long long)gm * d.A.ms + koff(d.A, gk);
#pragma unroll
        for (int p = 0; p < NPA; p++) As[p][r][kk] = ok ? ldg<T>(d.A.ptr, off + p * d.A.ps) : T(0);
      }
#pragma unroll
      for (int it = 0; it < (BK * BN) / NT; it++) {
        int e = tid + it * NT;
        int c, kk;
        if (b_nfast) { c = e % BN; kk = e / BN; } else { c = e / BK; kk = e % BK; }
        int gn = n0 + c, gk = k0 + kk;
        bool ok = gn < d.N && gk < d.K;
        long long off = bb + (long long)gn * d.B.ms + koff(d.B, gk);
#pragma unroll
        for (int p = 0; p < NPB; p++) Bs[p][kk][c] = ok ? ldg<T>(d.B.ptr, off + p * d.B.ps) : T(0);
      }
      __syncthreads();
#pragma unroll
      for (int ks = 0; ks < BK; ks += 4) {
        T a[2][NPA], bv[2][NPB];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
          for (int p = 0; p < NPA; p++) a[i][p] = As[p][wm * 32 + i * 16 + (lane & 15)][ks + (lane >> 4)];
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
          for (int p = 0; p < NPB; p++) bv[j][p] = Bs[p][ks + (lane >> 4)][wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
          for (int j = 0; j < 2; j++) {
            if constexpr (sizeof(T) == 8) {
              if constexpr (NPA == 2 && NPB == 2) {
                acc[i][j][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][0], bv[j][0], acc[i][j][0], 0, 0, 0);
                acc[i][j][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[i][1], bv[j][1], acc[i][j][0], 0, 0, 0);
                acc[i][j][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][0], bv[j][1], acc[i][j][1], 0, 0, 0);
                acc[i][j][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][1], bv[j][0], acc[i][j][1], 0, 0, 0);
              } else if constexpr (NPA == 2) {
                acc[i][j][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][0], bv[j][0], acc[i][j][0], 0, 0, 0);
                acc[i][j][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][1], bv[j][0], acc[i][j][1], 0, 0, 0);
              } else if constexpr (NPB == 2) {
                acc[i][j][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][0], bv[j][0], acc[i][j][0], 0, 0, 0);
                acc[i][j][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][0], bv[j][1], acc[i][j][1], 0, 0, 0);
              } else {
                acc[i][j][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][0], bv[j][0], acc[i][j][0], 0, 0, 0);
              }
            } else {
              if constexpr (NPA == 2 && NPB == 2) {
                acc[i][j][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][0], bv[j][0], acc[i][j][0], 0, 0, 0);
                acc[i][j][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(-a[i][1], bv[j][1], acc[i][j][0], 0, 0, 0);
                acc[i][j][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][0], bv[j][1], acc[i][j][1], 0, 0, 0);
                acc[i][j][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][1], bv[j][0], acc[i][j][1], 0, 0, 0);
              } else if constexpr (NPA == 2) {
                acc[i][j][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][0], bv[j][0], acc[i][j][0], 0, 0, 0);
                acc[i][j][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][1], bv[j][0], acc[i][j][1], 0, 0, 0);
              } else if constexpr (NPB == 2) {
                acc[i][j][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][0], bv[j][0], acc[i][j][0], 0, 0, 0);
                acc[i][j][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][0], bv[j][1], acc[i][j][1], 0, 0, 0);
              } else {
                acc[i][j][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][0], bv[j][0], acc[i][j][0], 0, 0, 0);
              }
            }
          }
      }
      __syncthreads();
    }
    // ---- epilogue.  Register q of a lane holds, in each 16x16 block,
    //   f32 16x16x4:  D[4*(lane/16) + q][lane%16]   (groups of 4 rows)
    //   f64 16x16x4:  D[4*q + lane/16][lane%16]     (4 groups of 1 row)
    const long long bc = (long long)(b / d.C.bdiv) * d.C.bs_hi + (long long)(b % d.C.bdiv) * d.C.bs_lo;
    T *C = reinterpret_cast<T *>(d.C.ptr);
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int rin = sizeof(T) == 8 ? 4 * q + (lane >> 4) : 4 * (lane >> 4) + q;
          int row = m0 + wm * 32 + i * 16 + rin;
          int col = n0 + wn * 32 + j * 16 + (lane & 15);
          if (row < d.M && col < d.N) {
            long long off = bc + (long long)row * d.C.rs + (long long)col * d.C.cs;
#pragma unroll
            for (int p = 0; p < NPC; p++) {
              T v = (T)d.alpha * acc[i][j][p][q];
              if (d.beta) v += C[off + p * d.C.ps];
              C[off + p * d.C.ps] = v;
            }
          }
        }
  }
}

// ---------------------------------------------------------------- bf16x3
// LDS holds hi and lo bf16 parts; A as [row][k], B transposed as [col][k],
// so that each lane's 8 consecutive K values are one 16-byte read.
template <int NPA, int NPB>
__global__ void __launch_bounds__(NT) kern_gemm_bf16x3(GemmDesc d) {
  constexpr int BK = 32, KP = BK + 8;  // padded row (80 bytes) keeps 16-byte alignment
  constexpr int NPC = (NPA > 1 || NPB > 1) ? 2 : 1;
  __shared__ __attribute__((aligned(16))) uint16_t Ah[NPA][BM][KP];
  __shared__ __attribute__((aligned(16))) uint16_t Al[NPA][BM][KP];
  __shared__ __attribute__((aligned(16))) uint16_t Bh[NPB][BN][KP];
  __shared__ __attribute__((aligned(16))) uint16_t Bl[NPB][BN][KP];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int tilesN = (d.N + BN - 1) / BN;
  const int m0 = (blockIdx.x / tilesN) * BM, n0 = (blockIdx.x % tilesN) * BN;
  const bool a_kfast = d.A.ks_lo == 1 && d.A.kdiv == 0;
  const bool b_nfast = d.B.ms == 1;

  for (int b = blockIdx.y; b < d.batch; b += gridDim.y) {
    const long long ba = boff(d.A, b), bb = boff(d.B, b);
    v4f acc[2][2][NPC];
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++)
        for (int p = 0; p < NPC; p++) acc[i][j][p] = v4f{0, 0, 0, 0};

    for (int k0 = 0; k0 < d.K; k0 += BK) {
#pragma unroll
      for (int it = 0; it < (BM * BK) / NT; it++) {
        int e = tid + it * NT;
        int r, kk;
        if (a_kfast) { r = e / BK; kk = e % BK; } else { r = e % BM; kk = e / BM; }
        int gm = m0 + r, gk = k0 + kk;
        bool ok = gm < d.M && gk < d.K;
        long long off = ba + (long long)gm * d.A.ms + koff(d.A, gk);
#pragma unroll
        for (int p = 0; p < NPA; p++) {
          float v = ok ? ldg<float>(d.A.ptr, off + p * d.A.ps) : 0.f;
          uint16_t h = f2bf(v);
          Ah[p][r][kk] = h;
          Al[p][r][kk] = f2bf(v - bf2f(h));
        }
      }
#pragma unroll
      for (int it = 0; it < (BK * BN) / NT; it++) {
        int e = tid + it * NT;
        int c, kk;
        if (b_nfast) { c = e % BN; kk = e / BN; } else { c = e / BK; kk = e % BK; }
        int gn = n0 + c, gk = k0 + kk;
        bool ok = gn < d.N && gk < d.K;
        long long off = bb + (long long)gn * d.B.ms + koff(d.B, gk);
#pragma unroll
        for (int p = 0; p < NPB; p++) {
          float v = ok ? ldg<float>(d.B.ptr, off + p * d.B.ps) : 0.f;
          uint16_t h = f2bf(v);
          Bh[p][c][kk] = h;
          Bl[p][c][kk] = f2bf(v - bf2f(h));
        }
      }
      __syncthreads();
      const int kq = 8 * (lane >> 4);
      uint4 ah[2][NPA], al[2][NPA], bh[2][NPB], bl[2][NPB];
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int p = 0; p < NPA; p++) {
          int r = wm * 32 + i * 16 + (lane & 15);
          ah[i][p] = *reinterpret_cast<const uint4 *>(&Ah[p][r][kq]);
          al[i][p] = *reinterpret_cast<const uint4 *>(&Al[p][r][kq]);
        }
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int p = 0; p < NPB; p++) {
          int c = wn * 32 + j * 16 + (lane & 15);
          bh[j][p] = *reinterpret_cast<const uint4 *>(&Bh[p][c][kq]);
          bl[j][p] = *reinterpret_cast<const uint4 *>(&Bl[p][c][kq]);
        }
      auto neg = [](uint4 v) { return uint4{v.x ^ 0x80008000u, v.y ^ 0x80008000u, v.z ^ 0x80008000u, v.w ^ 0x80008000u}; };
      auto mma3 = [](uint4 xh, uint4 xl, uint4 yh, uint4 yl, v4f c) {
        v8bf a_h = __builtin_bit_cast(v8bf, xh), a_l = __builtin_bit_cast(v8bf, xl);
        v8bf b_h = __builtin_bit_cast(v8bf, yh), b_l = __builtin_bit_cast(v8bf, yl);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_l, b_h, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h, b_l, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h, b_h, c, 0, 0, 0);
        return c;
      };
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
          if constexpr (NPA == 2 && NPB == 2) {
            acc[i][j][0] = mma3(ah[i][0], al[i][0], bh[j][0], bl[j][0], acc[i][j][0]);
            acc[i][j][0] = mma3(neg(ah[i][1]), neg(al[i][1]), bh[j][1], bl[j][1], acc[i][j][0]);
            acc[i][j][1] = mma3(ah[i][0], al[i][0], bh[j][1], bl[j][1], acc[i][j][1]);
            acc[i][j][1] = mma3(ah[i][1], al[i][1], bh[j][0], bl[j][0], acc[i][j][1]);
          } else if constexpr (NPA == 2) {
            acc[i][j][0] = mma3(ah[i][0], al[i][0], bh[j][0], bl[j][0], acc[i][j][0]);
            acc[i][j][1] = mma3(ah[i][1], al[i][1], bh[j][0], bl[j][0], acc[i][j][1]);
          } else if constexpr (NPB == 2) {
            acc[i][j][0] = mma3(ah[i][0], al[i][0], bh[j][0], bl[j][0], acc[i][j][0]);
            acc[i][j][1] = mma3(ah[i][0], al[i][0], bh[j][1], bl[j][1], acc[i][j][1]);
          } else {
            acc[i][j][0] = mma3(ah[i][0], al[i][0], bh[j][0], bl[j][0], acc[i][j][0]);
          }
        }
      __syncthreads();
    }
    const long long bc = (long long)(b / d.C.bdiv) * d.C.bs_hi + (long long)(b % d.C.bdiv) * d.C.bs_lo;
    float *C = reinterpret_cast<float *>(d.C.ptr);
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          int row = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + q;
          int col = n0 + wn * 32 + j * 16 + (lane & 15);
          if (row < d.M && col < d.N) {
            long long off = bc + (long long)row * d.C.rs + (long long)col * d.C.cs;
#pragma unroll
            for (int p = 0; p < NPC; p++) {
              float v = (float)d.alpha * acc[i][j][p][q];
              if (d.beta) v += C[off + p * d.C.ps];
              C[off + p * d.C.ps] = v;
            }
          }
        }
  }
}

template <typename K>
void launch(K kern, const GemmDesc &d, hipStream_t s) {
  int tiles = ((d.M + BM - 1) / BM) * ((d.N + BN - 1) / BN);
  int by = d.batch < 65535 ? d.batch : 65535;
  hipLaunchKernelGGL(kern, dim3(tiles, by), dim3(NT), 0, s, d);
}

}  // namespace

// ---------------------------------------------------------------- large real
// C = alpha A B (+ C) for one large real product in FP64 or FP32 MFMA: A (M x K)
// row-major with unit K stride, B (K x N) row-major with unit N stride.  128 x
// 128 tiles of four waves (64 x 64 each, 4 x 4 MFMA blocks of 16 x 16), K in
// slices of 16 with two LDS stages: the next slice is fetched into registers
// (two-element loads) while the current one is consumed, one barrier per
// slice.  Tiles are dealt in groups of 8 row tiles so that workgroups running
// together share the B tiles in L2.  K, N, the leading dimensions and the
// operand addresses must be even (element pairs never straddle an edge).
template <typename T>
__global__ void __launch_bounds__(NT) kern_gemm_big(const T *__restrict__ A, long long lda, const T *__restrict__ B,
                                                    long long ldb, T *C, long long ldc, int M, int N, int K, T alpha,
                                                    int beta) {
  constexpr int TM = 128, TN = 128, BK = 16;
  typedef T V2 __attribute__((ext_vector_type(2)));
  using V = typename std::conditional<sizeof(T) == 8, v4d, v4f>::type;
  __shared__ T As[2][TM][BK + 2];
  __shared__ T Bs[2][BK][TN + 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int tilesM = (M + TM - 1) / TM, tilesN = (N + TN - 1) / TN;
  const int per_group = 8 * tilesN, grp = blockIdx.x / per_group, first = grp * 8;
  const int gsz = min(8, tilesM - first), in = blockIdx.x % per_group;
  const int m0 = (first + in % gsz) * TM, n0 = (in / gsz) * TN;
  V acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = V{0, 0, 0, 0};
  V2 ra[4], rb[4];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int it = 0; it < 4; it++) {
      const int e = tid + NT * it;
      const int r = e >> 3, kp = (e & 7) * 2;
      const bool ok = m0 + r < M && k0 + kp < K;
      ra[it] = ok ? *reinterpret_cast<const V2 *>(A + (long long)(m0 + r) * lda + k0 + kp) : V2{0, 0};
      const int kk = e >> 6, np = (e & 63) * 2;
      const bool okb = k0 + kk < K && n0 + np < N;
      rb[it] = okb ? *reinterpret_cast<const V2 *>(B + (long long)(k0 + kk) * ldb + n0 + np) : V2{0, 0};
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int it = 0; it < 4; it++) {
      const int e = tid + NT * it;
      const int r = e >> 3, kp = (e & 7) * 2;
      As[buf][r][kp] = ra[it].x;
      As[buf][r][kp + 1] = ra[it].y;
      const int kk = e >> 6, np = (e & 63) * 2;
      *reinterpret_cast<V2 *>(&Bs[buf][kk][np]) = rb[it];
    }
  };
  const int nsl = (K + BK - 1) / BK;
  fetch(0);
  stage(0);
  __syncthreads();
  for (int s = 0; s < nsl; s++) {
    const int cur = s & 1;
    if (s + 1 < nsl) fetch((s + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      T a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++) a[i] = As[cur][wm * 64 + i * 16 + (lane & 15)][ks + (lane >> 4)];
#pragma unroll
      for (int j = 0; j < 4; j++) b[j] = Bs[cur][ks + (lane >> 4)][wn * 64 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
          if constexpr (sizeof(T) == 8) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
          else acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }
    if (s + 1 < nsl) stage(cur ^ 1);
    __syncthreads();
  }
  // register q of a lane: f64 D[4q + lane/16][lane%16], f32 D[4(lane/16) + q][lane%16]
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int rin = sizeof(T) == 8 ? 4 * q + (lane >> 4) : 4 * (lane >> 4) + q;
        const int row = m0 + wm * 64 + i * 16 + rin, col = n0 + wn * 64 + j * 16 + (lane & 15);
        if (row < M && col < N) {
          T *o = C + (long long)row * ldc + col;
          const T v = alpha * acc[i][j][q];
          *o = beta ? v + *o : v;
        }
      }
}

// the large-product kernel applies: one real product, unit K stride in A and
// unit N stride in B, even sizes and addresses, big enough to fill the chip
bool big_ok(const GemmDesc &d, size_t es) {
  const auto even = [es](const void *p) { return ((uintptr_t)p % (2 * es)) == 0; };
  return !d.ca && !d.cb && d.batch == 1 && d.A.ks_lo == 1 && d.A.kdiv == 0 && d.B.ms == 1 && d.B.kdiv == 0 &&
         d.C.cs == 1 && d.K % 2 == 0 && d.N % 2 == 0 && d.A.ms % 2 == 0 && d.B.ks_lo % 2 == 0 &&
         (long long)d.M * d.N >= 128LL * 128 * 64 && d.K >= 64 && even(d.A.ptr) && even(d.B.ptr);
}

template <typename T>
void launch_big(const GemmDesc &d, hipStream_t s) {
  const int tiles = ((d.M + 127) / 128) * ((d.N + 127) / 128);
  hipLaunchKernelGGL(kern_gemm_big<T>, dim3(tiles), dim3(NT), 0, s, (const T *)d.A.ptr, d.A.ms, (const T *)d.B.ptr,
                     d.B.ks_lo, (T *)d.C.ptr, d.C.rs, d.M, d.N, d.K, (T)d.alpha, d.beta);
}

void gemm(const GemmDesc &d, Prec prec, hipStream_t s) {
  if (d.M <= 0 || d.N <= 0 || d.batch <= 0) return;
  if (d.K <= 0) throw std::runtime_error("gemm: K must be positive");
  if (prec != PREC_BF16X3 && big_ok(d, prec == PREC_F64 ? 8 : 4) && !std::getenv("NED_NO_BIG_GEMM")) {
    if (prec == PREC_F64) launch_big<double>(d, s);
    else launch_big<float>(d, s);
    check_launch("gemm (large)");
    return;
  }
  const int key = (d.ca ? 2 : 0) + (d.cb ? 1 : 0);
  switch (prec) {
    case PREC_F64:
      if (key == 0) launch(kern_gemm_fp<double, 1, 1>, d, s);
      if (key == 1) launch(kern_gemm_fp<double, 1, 2>, d, s);
      if (key == 2) launch(kern_gemm_fp<double, 2, 1>, d, s);
      if (key == 3) launch(kern_gemm_fp<double, 2, 2>, d, s);
      break;
    case PREC_F32:
      if (key == 0) launch(kern_gemm_fp<float, 1, 1>, d, s);
      if (key == 1) launch(kern_gemm_fp<float, 1, 2>, d, s);
      if (key == 2) launch(kern_gemm_fp<float, 2, 1>, d, s);
      if (key == 3) launch(kern_gemm_fp<float, 2, 2>, d, s);
      break;
    case PREC_BF16X3:
      if (key == 0) launch(kern_gemm_bf16x3<1, 1>, d, s);
      if (key == 1) launch(kern_gemm_bf16x3<1, 2>, d, s);
      if (key == 2) launch(kern_gemm_bf16x3<2, 1>, d, s);
      if (key == 3) launch(kern_gemm_bf16x3<2, 2>, d, s);
      break;
  }
  check_launch("gemm");
}

}  // namespace ned

// ====================================================================
// Split-bf16 NT GEMM (the exchange and Pi products of the self-energy):
// C = alpha A B^T (+ C) with A (M x K) and B (N x K) both K-contiguous and
// pre-split into hi/lo bf16 planes.  Each product costs three
// v_mfma_f32_16x16x32_bf16 (lo*hi + hi*lo + hi*hi) with FP32 accumulation.
//
// The workgroup tile TM x TN is split into waves of WM x WN (each a grid of
// 16 x 16 MFMA blocks); the next 64-deep K slice is fetched into registers
// (16-byte loads) while the current one is consumed from LDS.  Tiles are
// visited in groups of 8 row-tiles so that concurrent workgroups share B
// tiles in L2.  Several tile shapes are compiled; the autotuner below times
// them on the first call for each problem shape and keeps the fastest.
namespace ned {
namespace {

// native 4 x 32-bit vector: keeps the prefetch registers out of scratch
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int NT_TK = 64;

typedef float v16f __attribute__((ext_vector_type(16)));

template <int TM, int TN, int WM, int WN, int TK = NT_TK, int MF = 16>
struct NtCfg {
  static constexpr int WAVES = (TM / WM) * (TN / WN);
  static constexpr int THREADS = 64 * WAVES;
  static constexpr int BI = WM / MF, BJ = WN / MF;  // MFMA blocks per wave
  static constexpr int SEG = TK / 8;                // 16-byte segments per row and K slice
  static constexpr int LA = TM * TK / 8 / THREADS, LB = TN * TK / 8 / THREADS;  // 16-byte loads per thread
};

// TK: depth of the K slice staged in LDS (rows padded by 8 elements, which
// keeps 16-byte ds_read_b128 conflict-free for 64 and 32).  STREAM: hold the
// B fragments of a 32-deep step and stream the A fragments one row block at
// a time (fewer live registers for tall wave tiles); the three products of a
// row block run as three passes over its column blocks, so consecutive MFMAs
// never share an accumulator.
// DB: two LDS stages; the next slice is written into the idle stage while the
// current one is consumed, one barrier per slice.
// IGLP >= 0: the compiler's MFMA/LDS-read interleaving strategy of that
// number (__builtin_amdgcn_iglp_opt; 0 interleaves small groups of fragment
// reads between the MFMAs), measured 2-3% faster on the nao = 64 shapes.
// SWZ (TK = 32 only): unpadded 64-byte LDS rows whose four 16-byte segments
// are permuted by row, segment s of row r stored at s ^ ((r >> 1) & 3).  A
// fragment read (16 rows, one segment per lane group) is then free of bank
// conflicts on gfx950, where the padded 80-byte rows measured a conflict
// cycle for every five active ones; a staging write (4 rows, all segments)
// stays one contiguous 256-byte run.
// MF: the matrix-core block, 16 (v_mfma_f32_16x16x32_bf16: a 16-row or
// 16-column operand fragment, 32 deep) or 32 (v_mfma_f32_32x32x16_bf16: 32
// rows or columns, 16 deep, twice the products per fragment read from LDS;
// lane l holds row / column l % 32 at depth 8 (l / 32) + 0..7, and output
// element e of lane l is row 8 (e / 4) + 4 (l / 32) + e % 4, column l % 32).
// DMA (two stages, swizzled): the next slice goes straight from memory into
// the idle LDS stage (global_load_lds_dwordx4: each wave instruction fills
// one contiguous 1 KB run, 16 rows of 64 bytes, lane l the 16 bytes at 16 l,
// so lane l fetches the logical segment that the swizzle stores there), with
// no register staging; the two stages are separate arrays, so that the
// compiler's wait for the transfers into one is not put before the reads of
// the other.
template <int TM, int TN, int WM, int WN, int TK = NT_TK, bool STREAM = false, bool DB = false, bool SWZ = false,
          int IGLP = -1, int MF = 16, bool DMA = false>
__global__ void __launch_bounds__(64 * (TM / WM) * (TN / WN))
    kern_nt_bf16x3(const uint16_t *__restrict__ Ahi, const uint16_t *__restrict__ Alo, long long lda, long long strA,
                   const uint16_t *__restrict__ Bhi, const uint16_t *__restrict__ Blo, long long ldb, long long strB,
                   float *__restrict__ C, long long ldc, long long strC, int M, int N, int K, float alpha, int beta) {
  using Cfg = NtCfg<TM, TN, WM, WN, TK, MF>;
  constexpr int NTH = Cfg::THREADS, BI = Cfg::BI, BJ = Cfg::BJ, LA = Cfg::LA, LB = Cfg::LB, SEG = Cfg::SEG;
  static_assert(!SWZ || TK == 32, "the swizzled layout is for 32-deep slices");
  static_assert(MF == 16 || (MF == 32 && !SWZ), "32 x 32 blocks read the padded layout");
  static_assert(!DMA || (DB && SWZ), "direct-to-LDS loads fill the swizzled two-stage layout");
  // one MFMA: depth KS, fragment rows / columns MF, lane groups of MF lanes
  constexpr int KS = MF == 16 ? 32 : 16;
  using Acc = typename std::conditional<MF == 16, v4f, v16f>::type;
  auto mfma = [](const v8bf &a, const v8bf &b, const Acc &c) __attribute__((always_inline)) {
    if constexpr (MF == 16) return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  };
  constexpr int LDK = SWZ ? TK : TK + 8;
  constexpr int WCOLS = TN / WN;
  // element offset of 8-element segment sg of LDS row rw
  auto lofs = [](int rw, int sg) __attribute__((always_inline)) {
    return SWZ ? ((sg ^ ((rw >> 1) & 3)) << 3) : (sg << 3);
  };
  // Workgroups are dealt round-robin to the 8 XCDs in launch order, each XCD
  // with its own L2: give every XCD a contiguous run of the (product, tile)
  // sequence -- every tile of a product of a batch on one XCD, at the same
  // time, so that they share its A and B rows in that L2 -- and, within a
  // product, of the group-ordered tile sequence, so that the 8-row x
  // few-column window an XCD works on at a time shares its tiles
  const int tilesN = N / TN, tilesM = M / TM, tiles = tilesM * tilesN;
  const int lin = blockIdx.y * gridDim.x + blockIdx.x, total = tiles * gridDim.y;
  const int xcd = lin % 8, perX = total / 8, rem = total % 8;
  const int unit = xcd * perX + min(xcd, rem) + lin / 8;
  const int item = unit / tiles, bid = unit % tiles;
  Ahi += (size_t)item * strA;
  Alo += (size_t)item * strA;
  Bhi += (size_t)item * strB;
  Blo += (size_t)item * strB;
  C += (size_t)item * strC;
  constexpr int NS = (DB && !DMA) ? 2 : 1;
  __shared__ __attribute__((aligned(16))) uint16_t sA[NS][2][TM][LDK];
  __shared__ __attribute__((aligned(16))) uint16_t sB[NS][2][TN][LDK];
  // the second stage of the DMA form (a placeholder otherwise)
  __shared__ __attribute__((aligned(16))) uint16_t sA1[DMA ? 2 : 1][DMA ? TM : 1][DMA ? LDK : 1];
  __shared__ __attribute__((aligned(16))) uint16_t sB1[DMA ? 2 : 1][DMA ? TN : 1][DMA ? LDK : 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WCOLS, wn = wave % WCOLS;
  const int group = 8;
  const int gsize = group * tilesN;
  const int g = bid / gsize, r = bid % gsize;
  const int gm = min(group, tilesM - g * group);
  const int tm = g * group + r % gm, tn = r / gm;
  const int m0 = tm * TM, n0 = tn * TN;

  Acc acc[BI][BJ];
#pragma unroll
  for (int i = 0; i < BI; i++)
#pragma unroll
    for (int j = 0; j < BJ; j++) acc[i][j] = Acc{};

  u32x4 ra0[LA], ra1[LA], rb0[LB], rb1[LB];
  auto gload = [&](int k0) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < LA; i++) {
      const int idx = tid + NTH * i, row = idx / SEG, seg = idx % SEG;
      const size_t o = (size_t)(m0 + row) * lda + k0 + seg * 8;
      ra0[i] = *reinterpret_cast<const u32x4 *>(Ahi + o);
      ra1[i] = *reinterpret_cast<const u32x4 *>(Alo + o);
    }
#pragma unroll
    for (int i = 0; i < LB; i++) {
      const int idx = tid + NTH * i, row = idx / SEG, seg = idx % SEG;
      const size_t o = (size_t)(n0 + row) * ldb + k0 + seg * 8;
      rb0[i] = *reinterpret_cast<const u32x4 *>(Bhi + o);
      rb1[i] = *reinterpret_cast<const u32x4 *>(Blo + o);
    }
  };
  auto lstore = [&](int st) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < LA; i++) {
      const int idx = tid + NTH * i, row = idx / SEG, seg = idx % SEG;
      *reinterpret_cast<u32x4 *>(&sA[st][0][row][lofs(row, seg)]) = ra0[i];
      *reinterpret_cast<u32x4 *>(&sA[st][1][row][lofs(row, seg)]) = ra1[i];
    }
#pragma unroll
    for (int i = 0; i < LB; i++) {
      const int idx = tid + NTH * i, row = idx / SEG, seg = idx % SEG;
      *reinterpret_cast<u32x4 *>(&sB[st][0][row][lofs(row, seg)]) = rb0[i];
      *reinterpret_cast<u32x4 *>(&sB[st][1][row][lofs(row, seg)]) = rb1[i];
    }
  };
  auto compute = [&](const uint16_t(*cA)[TM][LDK], const uint16_t(*cB)[TN][LDK]) __attribute__((always_inline)) {
    if constexpr (IGLP >= 0) __builtin_amdgcn_iglp_opt(IGLP);
#pragma unroll
    for (int kk = 0; kk < TK; kk += KS) {
      const int kseg = (kk >> 3) + lane / MF;
      if constexpr (STREAM) {
        v8bf bh[BJ], bl[BJ];
#pragma unroll
        for (int j = 0; j < BJ; j++) {
          const int col = wn * WN + j * MF + lane % MF;
          bh[j] = __builtin_bit_cast(v8bf, *reinterpret_cast<const u32x4 *>(&cB[0][col][lofs(col, kseg)]));
          bl[j] = __builtin_bit_cast(v8bf, *reinterpret_cast<const u32x4 *>(&cB[1][col][lofs(col, kseg)]));
        }
#pragma unroll
        for (int i = 0; i < BI; i++) {
          const int row = wm * WM + i * MF + lane % MF;
          const v8bf a_h = __builtin_bit_cast(v8bf, *reinterpret_cast<const u32x4 *>(&cA[0][row][lofs(row, kseg)]));
          const v8bf a_l = __builtin_bit_cast(v8bf, *reinterpret_cast<const u32x4 *>(&cA[1][row][lofs(row, kseg)]));
#pragma unroll
          for (int j = 0; j < BJ; j++) acc[i][j] = mfma(a_l, bh[j], acc[i][j]);
#pragma unroll
          for (int j = 0; j < BJ; j++) acc[i][j] = mfma(a_h, bl[j], acc[i][j]);
#pragma unroll
          for (int j = 0; j < BJ; j++) acc[i][j] = mfma(a_h, bh[j], acc[i][j]);
        }
      } else {
        u32x4 ah[BI], al[BI], bh[BJ], bl[BJ];
#pragma unroll
        for (int i = 0; i < BI; i++) {
          const int row = wm * WM + i * MF + lane % MF;
          ah[i] = *reinterpret_cast<const u32x4 *>(&cA[0][row][lofs(row, kseg)]);
          al[i] = *reinterpret_cast<const u32x4 *>(&cA[1][row][lofs(row, kseg)]);
        }
#pragma unroll
        for (int j = 0; j < BJ; j++) {
          const int col = wn * WN + j * MF + lane % MF;
          bh[j] = *reinterpret_cast<const u32x4 *>(&cB[0][col][lofs(col, kseg)]);
          bl[j] = *reinterpret_cast<const u32x4 *>(&cB[1][col][lofs(col, kseg)]);
        }
        // three passes over every (i, j): BI x BJ independent accumulators lie
        // between two products into the same one (the 32 x 32 MFMA's result
        // latency is twice the 16 x 16 one's)
#pragma unroll
        for (int pass = 0; pass < 3; pass++)
#pragma unroll
          for (int i = 0; i < BI; i++)
#pragma unroll
            for (int j = 0; j < BJ; j++) {
              const v8bf a = __builtin_bit_cast(v8bf, pass == 0 ? al[i] : ah[i]);
              const v8bf b = __builtin_bit_cast(v8bf, pass == 1 ? bl[j] : bh[j]);
              acc[i][j] = mfma(a, b, acc[i][j]);
            }
      }
    }
  };
  if constexpr (DMA) {
    // 1 KB runs of a slice (16 rows x 64 bytes of one plane): A's planes,
    // then B's; run WV i + w goes to wave w, so that the operand and plane of
    // a wave's i-th run are known at compile time
    constexpr int WV = NTH / 64, RA = 2 * (TM / 16), RUNS = RA + 2 * (TN / 16);
    static_assert((TM / 16) % WV == 0 && (TN / 16) % WV == 0, "a plane's runs in whole rounds of the waves");
    const uint16_t *const gAh = Ahi, *const gAl = Alo, *const gBh = Bhi, *const gBl = Blo;
    auto dma = [&](uint16_t(*dA)[TM][LDK], uint16_t(*dB)[TN][LDK], int k0) __attribute__((always_inline)) {
#pragma unroll
      for (int i = 0; i < RUNS / WV; i++) {
        const int q0 = WV * i;
        const bool isA = q0 < RA;
        const int per = (isA ? TM : TN) / 16, qq0 = isA ? q0 : q0 - RA;
        const int pl = qq0 / per, rb = qq0 % per + wave;
        const int rw = 16 * rb + (lane >> 2), sg = (lane & 3) ^ ((rw >> 1) & 3);
        const uint16_t *src = isA ? (pl ? gAl : gAh) + (size_t)(m0 + rw) * lda + k0 + 8 * sg
                                  : (pl ? gBl : gBh) + (size_t)(n0 + rw) * ldb + k0 + 8 * sg;
        uint16_t *dst = isA ? &dA[pl][16 * rb][0] : &dB[pl][16 * rb][0];
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)src,
                                         (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
      }
    };
    dma(sA[0], sB[0], 0);
    __syncthreads();  // waits for this wave's transfers, then for every wave
    for (int k0 = 0; k0 < K; k0 += 2 * TK) {
      if (k0 + TK < K) dma(sA1, sB1, k0 + TK);
      compute(sA[0], sB[0]);
      __syncthreads();
      if (k0 + TK < K) {
        if (k0 + 2 * TK < K) dma(sA[0], sB[0], k0 + 2 * TK);
        compute(sA1, sB1);
        __syncthreads();
      }
    }
  } else if constexpr (DB) {
    gload(0);
    lstore(0);
    if (TK < K) gload(TK);
    __syncthreads();
    int st = 0;
    for (int k0 = 0; k0 < K; k0 += TK) {
      if (k0 + TK < K) {
        lstore(st ^ 1);
        if (k0 + 2 * TK < K) gload(k0 + 2 * TK);
      }
      compute(sA[st], sB[st]);
      __syncthreads();
      st ^= 1;
    }
  } else {
    gload(0);
    for (int k0 = 0; k0 < K; k0 += TK) {
      __syncthreads();
      lstore(0);
      __syncthreads();
      if (k0 + TK < K) gload(k0 + TK);
      compute(sA[0], sB[0]);
    }
  }
#pragma unroll
  for (int i = 0; i < BI; i++)
#pragma unroll
    for (int j = 0; j < BJ; j++)
#pragma unroll
      for (int q = 0; q < (MF == 16 ? 4 : 16); q++) {
        const int row = m0 + wm * WM + i * MF + (MF == 16 ? 4 * (lane >> 4) + q : 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3));
        const int col = n0 + wn * WN + j * MF + lane % MF;
        float *cp = C + (size_t)row * ldc + col;
        float v = alpha * acc[i][j][q];
        if (beta) v += *cp;
        *cp = v;
      }
}

__global__ void kern_split_bf16(const float *__restrict__ x, long long ldx, uint16_t *__restrict__ hi,
                                uint16_t *__restrict__ lo, long long ldh, long long rows, int cols) {
  const int c4 = cols >> 2;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * c4) return;
  const long long r = i / c4;
  const int c = (int)(i % c4) * 4;
  const float4 v = *reinterpret_cast<const float4 *>(x + r * ldx + c);
  const float f[4] = {v.x, v.y, v.z, v.w};
  uint16_t h[4], l[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    h[q] = f2bf(f[q]);
    l[q] = f2bf(f[q] - bf2f(h[q]));
  }
  *reinterpret_cast<uint2 *>(hi + r * ldh + c) =
      uint2{(uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16)};
  *reinterpret_cast<uint2 *>(lo + r * ldh + c) =
      uint2{(uint32_t)l[0] | ((uint32_t)l[1] << 16), (uint32_t)l[2] | ((uint32_t)l[3] << 16)};
}

// ---- tile shapes and the autotuner
struct NtVariant {
  int tm, tn;
  void (*launch)(const NtOp &, const NtOp &, float *, long long, long long, int, int, int, int, float, int,
                 hipStream_t);
  const char *name;
};

template <int TM, int TN, int WM, int WN, int TK = NT_TK, bool STREAM = false, bool DB = false, bool SWZ = false,
          int IGLP = -1, int MF = 16, bool DMA = false>
void nt_launch(const NtOp &A, const NtOp &B, float *C, long long ldc, long long strC, int M, int N, int K, int batch,
               float alpha, int beta, hipStream_t s) {
  // the grid covers whole tiles only: a ragged M or N would leave rows or
  // columns unwritten
  if (M % TM || N % TN) throw std::invalid_argument("gemm_nt_bf16x3: tile does not divide M or N");
  const int tiles = (M / TM) * (N / TN);
  hipLaunchKernelGGL((kern_nt_bf16x3<TM, TN, WM, WN, TK, STREAM, DB, SWZ, IGLP, MF, DMA>), dim3(tiles, batch),
                     dim3(NtCfg<TM, TN, WM, WN, TK, MF>::THREADS), 0, s, A.hi, A.lo, A.ld, A.stride, B.hi, B.lo, B.ld,
                     B.stride, C, ldc, strC, M, N, K, alpha, beta);
}

const NtVariant kNt[] = {
    {128, 128, nt_launch<128, 128, 64, 64>, "128x128 / 4 waves of 64x64"},
    {256, 128, nt_launch<256, 128, 128, 64>, "256x128 / 4 waves of 128x64"},
    {128, 256, nt_launch<128, 256, 64, 128>, "128x256 / 4 waves of 64x128"},
    {256, 128, nt_launch<256, 128, 64, 64>, "256x128 / 8 waves of 64x64"},
    {128, 128, nt_launch<128, 128, 32, 64>, "128x128 / 8 waves of 32x64"},
    {256, 256, nt_launch<256, 256, 128, 64, 32, true>, "256x256 / 8 waves of 128x64, K32, streamed A"},
    {128, 128, nt_launch<128, 128, 64, 64, 64, true>, "128x128 / 4 waves of 64x64, streamed A"},
    {256, 256, nt_launch<256, 256, 128, 64, 32, true, true>, "256x256 / 8 waves of 128x64, K32, streamed A, 2 stages"},
    {128, 128, nt_launch<128, 128, 64, 64, 32, true, true>, "128x128 / 4 waves of 64x64, K32, streamed A, 2 stages"},
    {256, 128, nt_launch<256, 128, 64, 64, 32, true, true>, "256x128 / 8 waves of 64x64, K32, streamed A, 2 stages"},
    {256, 256, nt_launch<256, 256, 128, 64, 32, true, true, true>,
     "256x256 / 8 waves of 128x64, K32, streamed A, 2 stages, swizzled"},
    {128, 128, nt_launch<128, 128, 64, 64, 32, true, true, true>,
     "128x128 / 4 waves of 64x64, K32, streamed A, 2 stages, swizzled"},
    {256, 128, nt_launch<256, 128, 64, 64, 32, true, true, true>,
     "256x128 / 8 waves of 64x64, K32, streamed A, 2 stages, swizzled"},
    {256, 256, nt_launch<256, 256, 128, 64, 32, true, true, true, 0>,
     "256x256 / 8 waves of 128x64, K32, streamed A, 2 stages, swizzled, MFMA/LDS interleave"},
    {256, 128, nt_launch<256, 128, 64, 64, 64, false, false, false, 0>,
     "256x128 / 8 waves of 64x64, MFMA/LDS interleave"},
    {256, 256, nt_launch<256, 256, 128, 64, 32, true, true, false, 0, 32>,
     "256x256 / 8 waves of 128x64, K32, streamed A, 2 stages, MFMA/LDS interleave, 32x32 blocks"},
    {256, 256, nt_launch<256, 256, 128, 64, 32, false, true, false, 0, 32>,
     "256x256 / 8 waves of 128x64, K32, 2 stages, MFMA/LDS interleave, 32x32 blocks"},
    {128, 128, nt_launch<128, 128, 64, 64, 32, true, true, false, 0, 32>,
     "128x128 / 4 waves of 64x64, K32, streamed A, 2 stages, MFMA/LDS interleave, 32x32 blocks"},
    {256, 128, nt_launch<256, 128, 64, 64, 32, true, true, false, 0, 32>,
     "256x128 / 8 waves of 64x64, K32, streamed A, 2 stages, MFMA/LDS interleave, 32x32 blocks"},
    {256, 256, nt_launch<256, 256, 128, 64, 32, true, true, true, -1, 16, true>,
     "256x256 / 8 waves of 128x64, K32, streamed A, 2 stages, swizzled, direct-to-LDS loads"},
    {128, 128, nt_launch<128, 128, 64, 64, 32, true, true, true, -1, 16, true>,
     "128x128 / 4 waves of 64x64, K32, streamed A, 2 stages, swizzled, direct-to-LDS loads"},
};
constexpr int kNtCount = sizeof(kNt) / sizeof(kNt[0]);

struct NtKey {
  int M, N, K, batch, mfit;
  bool operator<(const NtKey &o) const {
    return std::tie(M, N, K, batch, mfit) < std::tie(o.M, o.N, o.K, o.batch, o.mfit);
  }
};
std::mutex g_nt_mu;
std::map<std::pair<int, NtKey>, int> g_nt_choice;  // (device, shape) -> variant

// NED_NT_VARIANT, read on every call so that a process can pin the variants
// one after another
int nt_forced() {
  const char *e = std::getenv("NED_NT_VARIANT");
  const int v = e ? std::atoi(e) : -1;
  return (v >= 0 && v < kNtCount) ? v : -1;
}

bool fits(const NtVariant &v, int M, int N) { return M % v.tm == 0 && N % v.tn == 0; }

}  // namespace

bool nt_bf16x3_ok(int M, int N, int K) { return M % 128 == 0 && N % 128 == 0 && K % NT_TK == 0; }

int nt_variant_count() { return kNtCount; }
const char *nt_variant_name(int v) { return v >= 0 && v < kNtCount ? kNt[v].name : "?"; }

// The autotuner: on the first call for a (device, M, N, K, batch) the fitting
// variants run on the caller's stream (C is overwritten by the timing runs
// and then by the real one) and the fastest is kept for later calls.
int nt_choose(const NtOp &A, const NtOp &B, float *C, long long ldc, long long strC, int M, int N, int K, int batch,
              hipStream_t s, bool tune) {
  const int forced = nt_forced();
  if (forced >= 0 && fits(kNt[forced], M, N)) return forced;
  int dev = 0;
  NED_HIP(hipGetDevice(&dev));
  // M and batch enter by octave, so the smaller last chunk of a slice reuses
  // the choice made for the full chunks
  auto octave = [](int x) {
    int o = 0;
    while (x > 1) { x >>= 1; o++; }
    return o;
  };
  // ... and the largest row tile that divides M is part of the key: an M in
  // the same octave that only 128-row tiles divide must not reuse a 256-row
  // choice
  const int mfit = M % 256 == 0 ? 256 : 128;
  const std::pair<int, NtKey> key{dev, NtKey{octave(M), N, K, octave(batch), mfit}};
  {
    std::lock_guard<std::mutex> lk(g_nt_mu);
    auto it = g_nt_choice.find(key);
    if (it != g_nt_choice.end() && fits(kNt[it->second], M, N)) return it->second;
  }
  if (!tune) {
    // untimed default: the streamed, double-buffered, swizzled tiles, the
    // largest that fits
    for (int v : {10, 12, 11})
      if (fits(kNt[v], M, N)) return v;
    return 0;
  }
  hipEvent_t e0, e1;
  NED_HIP(hipEventCreate(&e0));
  NED_HIP(hipEventCreate(&e1));
  int best = 0;
  float best_ms = 1e30f;
  for (int v = 0; v < kNtCount; v++) {
    if (!fits(kNt[v], M, N)) continue;
    kNt[v].launch(A, B, C, ldc, strC, M, N, K, batch, 1.0f, 0, s);  // warm-up
    // the best of three timed launches: variants a few percent apart would
    // otherwise trade places from run to run
    float ms = 1e30f;
    for (int rep = 0; rep < 3; rep++) {
      NED_HIP(hipEventRecord(e0, s));
      kNt[v].launch(A, B, C, ldc, strC, M, N, K, batch, 1.0f, 0, s);
      NED_HIP(hipEventRecord(e1, s));
      NED_HIP(hipEventSynchronize(e1));
      check_launch("gemm_nt_bf16x3 autotune");
      float t = 0;
      NED_HIP(hipEventElapsedTime(&t, e0, e1));
      ms = std::min(ms, t);
    }
    if (std::getenv("NED_NT_VERBOSE"))
      std::fprintf(stderr, "[nt autotune] M=%d N=%d K=%d batch=%d  %-28s %.3f ms\n", M, N, K, batch, kNt[v].name, ms);
    if (ms < best_ms) {
      best_ms = ms;
      best = v;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::lock_guard<std::mutex> lk(g_nt_mu);
  g_nt_choice[key] = best;
  return best;
}

void gemm_nt_bf16x3(const NtOp &A, const NtOp &B, float *C, long long ldc, long long strC, int M, int N, int K,
                    int batch, float alpha, int beta, hipStream_t s, bool tune) {
  if (!nt_bf16x3_ok(M, N, K) || A.ld % 8 || B.ld % 8 || A.stride % 8 || B.stride % 8)
    throw std::invalid_argument("gemm_nt_bf16x3: M, N must be multiples of 128, K of 64, strides of 8");
  // beta = 1 reads C, which the timing runs would overwrite: autotune only on beta = 0
  const int v = beta ? 0 : nt_choose(A, B, C, ldc, strC, M, N, K, batch, s, tune);
  kNt[v].launch(A, B, C, ldc, strC, M, N, K, batch, alpha, beta, s);
  check_launch("gemm_nt_bf16x3");
}

void split_bf16(const float *x, long long ldx, uint16_t *hi, uint16_t *lo, long long ldh, long long rows, int cols,
                hipStream_t s) {
  if (cols % 4 || ldx % 4 || ldh % 4) throw std::invalid_argument("split_bf16: cols and strides must be multiples of 4");
  const long long th = rows * (cols / 4);
  if (th == 0) return;
  hipLaunchKernelGGL(kern_split_bf16, dim3((unsigned)((th + 255) / 256)), dim3(256), 0, s, x, ldx, hi, lo, ldh, rows,
                     cols);
  check_launch("split_bf16");
}

}  // namespace ned
