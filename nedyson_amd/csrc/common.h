// Shared host/device definitions for the nedyson_amd HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace ned {

// Arithmetic used by the Sigma contractions.
enum Prec : int { PREC_F64 = 0, PREC_F32 = 1, PREC_BF16X3 = 2 };

#define NED_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) +     \
                               " at " + __FILE__ + ":" + std::to_string(__LINE__));     \
  } while (0)

inline void check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw std::runtime_error(std::string("kernel launch failed (") + what + "): " + hipGetErrorString(e));
}

// complex FP64 on device: interleaved (re, im) pairs, the layout of
// std::complex<double> used by green_func.
struct alignas(16) z2 {
  double x, y;
};

__host__ __device__ inline z2 zmul(z2 a, z2 b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__host__ __device__ inline z2 zadd(z2 a, z2 b) { return {a.x + b.x, a.y + b.y}; }
__host__ __device__ inline z2 zsub(z2 a, z2 b) { return {a.x - b.x, a.y - b.y}; }
__host__ __device__ inline z2 zconj(z2 a) { return {a.x, -a.y}; }
__host__ __device__ inline z2 zscale(z2 a, double s) { return {a.x * s, a.y * s}; }

// one element of a complex history in FP64 (z2) or FP32 (float2) storage:
// ldz widens to FP64, stz rounds to the storage type (to nearest)
__device__ __forceinline__ z2 ldz(const z2 *p) { return *p; }
__device__ __forceinline__ z2 ldz(const float2 *p) {
  const float2 v = *p;
  return z2{(double)v.x, (double)v.y};
}
__device__ __forceinline__ void stz(z2 *p, z2 v) { *p = v; }
__device__ __forceinline__ void stz(float2 *p, z2 v) { *p = float2{(float)v.x, (float)v.y}; }

// bf16 round-to-nearest-even of an FP32 value (finite inputs), and back
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

// ------------------------------------------------------------------ GEMM
// One operand of the strided batched GEMM.  The logical matrix is A (M x K)
// or B (K x N); element (r, c) of batch b sits at
//   ptr + boff(b) + r*rs + koff(c or r)   (see kern_gemm)
// where the K index may be split in two levels, kk = q*kdiv + r, with
// offset q*ks_hi + r*ks_lo (kdiv == 0: plain stride ks_lo).  Complex
// operands are planar: the imaginary plane sits ps elements after the real.
struct GOp {
  const void *ptr = nullptr;
  long long ms = 0;        // stride of the non-K index (row of A / column of B)
  long long ks_lo = 0;     // stride of the K index (low level)
  long long ks_hi = 0;     // stride of the high level of the K index
  int kdiv = 0;            // 0: single-level K
  long long ps = 0;        // plane stride (complex)
  int bdiv = 1 << 30;      // batch b -> (b / bdiv) * bs_hi + (b % bdiv) * bs_lo
  long long bs_hi = 0, bs_lo = 0;  // (single-level batches use bs_lo)
};

struct GOut {
  void *ptr = nullptr;
  long long rs = 0, cs = 1, ps = 0;
  int bdiv = 1 << 30;
  long long bs_hi = 0, bs_lo = 0;
};

struct GemmDesc {
  int M = 0, N = 0, K = 0, batch = 1;
  bool ca = false, cb = false;  // complex operands
  GOp A, B;
  GOut C;
  double alpha = 1.0;
  int beta = 0;  // 0: overwrite, 1: accumulate
};

// Launch C = alpha * A B (+ C) in the arithmetic of prec on stream s.
//   PREC_F64: operands/result double, v_mfma_f64_16x16x4f64
//   PREC_F32: operands/result float,  v_mfma_f32_16x16x4f32
//   PREC_BF16X3: operands/result float, each operand split a = hi + lo in
//                bf16, product hi*hi + hi*lo + lo*hi on
//                v_mfma_f32_16x16x32_bf16 with FP32 accumulation
void gemm(const GemmDesc &d, Prec prec, hipStream_t s);

// Split-bf16 NT GEMM for the self-energy exchange contraction:
// C (M x N, row stride ldc) = alpha * A B^T (+ C if beta), A = Ahi + Alo
// (M x K), B = Bhi + Blo (N x K), K-contiguous bf16 planes.
// Operand: hi/lo planes, row stride ld (elements), batch stride.
struct NtOp {
  const uint16_t *hi, *lo;
  long long ld, stride;
};
bool nt_bf16x3_ok(int M, int N, int K);
// tile-shape variants of gemm_nt_bf16x3 (the autotuner's candidates; the
// environment variable NED_NT_VARIANT=<index> pins one)
int nt_variant_count();
const char *nt_variant_name(int v);
// tune = false: no timing runs (the tile variant chosen earlier for this
// shape, else a fixed default)
void gemm_nt_bf16x3(const NtOp &A, const NtOp &B, float *C, long long ldc, long long sC, int M, int N, int K,
                    int batch, float alpha, int beta, hipStream_t s, bool tune = true);
// x -> hi + lo with hi = bf16(x), lo = bf16(x - hi), round to nearest even;
// rows x cols (cols a multiple of 4) from row stride ldx to row stride ldh
void split_bf16(const float *x, long long ldx, uint16_t *hi, uint16_t *lo, long long ldh, long long rows, int cols,
                hipStream_t s);

}  // namespace ned
