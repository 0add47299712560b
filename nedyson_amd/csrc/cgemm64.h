// Batched 64 x 64 complex products of the self-energy chain (split-bf16,
// nao = 64); see cgemm64.hip for the operand layouts.  G = auxiliary indices
// per workgroup (S = number of partial sums for c64_e).
#pragma once
#include "common.h"

namespace ned {

// W == nullptr: only the transposed split Wth / Wtl is written
void c64_w(const uint16_t *Vh, const uint16_t *Vl, const float *At, int P, int a, int G, float *W, uint16_t *Wth,
           uint16_t *Wtl, hipStream_t s);
void c64_z(const uint16_t *Vh, const uint16_t *Vl, const float *Bt, int P, int a, int G, uint16_t *Zth,
           uint16_t *Ztl, hipStream_t s);
// pairs != nullptr: the row in the split layout of the exchange (pair table,
// antisymmetric block from column ks); otherwise the 64 x 64 block row-major
void c64_c(const float *Bm, const uint16_t *Wth, const uint16_t *Wtl, int P, int a, int G, uint16_t *Sh,
           uint16_t *Sl, long long ld, hipStream_t s, const uint16_t *pairs = nullptr, long long ks = 0);
// Y != nullptr: H^a from the split exchange's outputs (row stride ldy, the
// antisymmetric part from column nsp) instead of Ht, and W^a from its
// transposed split Wth / Wtl (as c64_w writes it) instead of W
void c64_e(const float *W, const float *Ht, int P, int a, int S, float *Ep, hipStream_t s, const float *Y = nullptr,
           long long ldy = 0, int nsp = 0, const uint16_t *pairs = nullptr, const uint16_t *Wth = nullptr,
           const uint16_t *Wtl = nullptr);

}  // namespace ned
