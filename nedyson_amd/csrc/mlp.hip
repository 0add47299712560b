// The learned next-slice predictor: a three-layer perceptron
//   x (n_in) -> tanh(W1 x + b1) -> tanh(W2 h1 + b2) -> W3 h2 + b3 (n_out)
// trained and evaluated on the GPU.  Training's matrix products go through
// this library's own MFMA GEMM (gemm(), FP32); inference can instead run with
// bf16 weights and activations on the bf16 matrix cores (forward_bf16, FP32
// accumulation).  The rest are the element-wise, reduction and Adam kernels
// below.  Rows of a batch are independent.
//
// The network has no counterpart in the reference: it replaces the
// polynomial predictor dyson::Extrapolate (src/dyson.cpp:47-117) as the
// corrector's initial guess on a few time points of each slice (see
// predictor.hip), and the plain iteration still decides convergence.
#include "mlp.h"
#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

namespace ned {
namespace {

__global__ void kern_bias_act(float *Z, const float *b, long long rows, int cols, int act) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * cols) return;
  float v = Z[i] + b[i % cols];
  Z[i] = act ? tanhf(v) : v;
}

// dY = 2 (Y - T) / count; partial sums of (Y - T)^2 per workgroup
__global__ void kern_mse(const float *Y, const float *T, long long n, float scale, float *dY, double *part) {
  __shared__ double red[256];
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0;
  if (i < n) {
    const float d = Y[i] - T[i];
    dY[i] = 2.0f * d * scale;
    s = (double)d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void kern_sum_d(const double *part, int n, double *out) {
  __shared__ double red[256];
  double s = 0;
  for (int i = threadIdx.x; i < n; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// dZ = dH * (1 - H^2)  (H = tanh(Z))
__global__ void kern_tanh_back(float *dH, const float *H, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dH[i] *= 1.0f - H[i] * H[i];
}

// db[c] = sum_r dZ[r][c]
__global__ void kern_colsum(const float *dZ, long long rows, int cols, float *db) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  float s = 0;
  for (long long r = 0; r < rows; r++) s += dZ[r * cols + c];
  db[c] = s;
}

// Adam (Kingma & Ba 2015, Algorithm 1) with bias-corrected moments
__global__ void kern_adam(float *p, const float *g, float *m, float *v, long long n, float lr, float b1, float b2,
                          float eps, float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1 - b1) * gi;
  const float vi = b2 * v[i] + (1 - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// hi[r][c] = bf16(x[r][c]) (round to nearest even) for r < rows, 0 for the
// padding rows up to rows_pad
__global__ void kern_to_bf16(const float *x, long long rows, long long rows_pad, int cols, uint16_t *hi) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows_pad * cols) return;
  hi[i] = i / cols < rows ? f2bf(x[i]) : (uint16_t)0;
}

// the next layer's bf16 input from a layer's product o: bf16(tanh(o + b)) for
// r < rows, 0 for the padding rows up to rows_pad
__global__ void kern_bias_tanh_bf16(const float *o, const float *b, long long rows, long long rows_pad, int cols,
                                    uint16_t *out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows_pad * cols) return;
  out[i] = i / cols < rows ? f2bf(tanhf(o[i] + b[i % cols])) : (uint16_t)0;
}

// y[r][c] = o[r][c] + b[c] for r < rows
__global__ void kern_bias_rows(const float *o, const float *b, long long rows, int cols, float *y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * cols) return;
  y[i] = o[i] + b[i % cols];
}

__global__ void kern_gather_rows(const float *src, const int *idx, int rows, int cols, float *dst) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)rows * cols) return;
  const int r = (int)(i / cols), c = (int)(i % cols);
  dst[i] = src[(long long)idx[r] * cols + c];
}

inline unsigned nb(long long n) { return (unsigned)((n + 255) / 256); }

// C (M x N, row-major) = A (M x K, row-major) B^T  with B (N x K) row-major
void mm_nt(const float *A, const float *B, float *C, int M, int N, int K, hipStream_t s) {
  GemmDesc d;
  d.M = M; d.N = N; d.K = K;
  d.A.ptr = A; d.A.ms = K; d.A.ks_lo = 1;
  d.B.ptr = B; d.B.ms = K; d.B.ks_lo = 1;
  d.C.ptr = C; d.C.rs = N; d.C.cs = 1;
  gemm(d, PREC_F32, s);
}
// C (M x N) = A^T B  with A (K x M), B (K x N) row-major
void mm_tn(const float *A, const float *B, float *C, int M, int N, int K, hipStream_t s) {
  GemmDesc d;
  d.M = M; d.N = N; d.K = K;
  d.A.ptr = A; d.A.ms = 1; d.A.ks_lo = M;
  d.B.ptr = B; d.B.ms = 1; d.B.ks_lo = N;
  d.C.ptr = C; d.C.rs = N; d.C.cs = 1;
  gemm(d, PREC_F32, s);
}
// C (M x N) = A B  with A (M x K), B (K x N) row-major
void mm_nn(const float *A, const float *B, float *C, int M, int N, int K, hipStream_t s) {
  GemmDesc d;
  d.M = M; d.N = N; d.K = K;
  d.A.ptr = A; d.A.ms = K; d.A.ks_lo = 1;
  d.B.ptr = B; d.B.ms = 1; d.B.ks_lo = N;
  d.C.ptr = C; d.C.rs = N; d.C.cs = 1;
  gemm(d, PREC_F32, s);
}

}  // namespace

Mlp::Mlp(int in, int width, int out, int device_, unsigned long long seed) : n_in(in), w(width), n_out(out), device(device_) {
  NED_HIP(hipSetDevice(device));
  NED_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  off[0] = 0;
  off[1] = off[0] + (long long)w * n_in;
  off[2] = off[1] + w;
  off[3] = off[2] + (long long)w * w;
  off[4] = off[3] + w;
  off[5] = off[4] + (long long)n_out * w;
  off[6] = off[5] + n_out;
  nparam = off[6];
  params.ensure(nparam * sizeof(float));
  grads.ensure(nparam * sizeof(float));
  adam_m.ensure(nparam * sizeof(float));
  adam_v.ensure(nparam * sizeof(float));
  // Glorot-uniform weights, zero biases
  std::mt19937_64 rng(seed);
  std::vector<float> h(nparam, 0.0f);
  auto fill = [&](long long o, int fo, int fi) {
    const float lim = std::sqrt(6.0f / (float)(fi + fo));
    std::uniform_real_distribution<float> U(-lim, lim);
    for (long long i = 0; i < (long long)fo * fi; i++) h[o + i] = U(rng);
  };
  fill(off[0], w, n_in);
  fill(off[2], w, w);
  fill(off[4], n_out, w);
  NED_HIP(hipMemcpy(params.p, h.data(), nparam * sizeof(float), hipMemcpyHostToDevice));
  reset_adam();
}

Mlp::~Mlp() {
  if (stream) (void)hipStreamDestroy(stream);
}

void Mlp::reset_adam() {
  std::vector<float> z(nparam, 0.0f);
  NED_HIP(hipMemcpy(adam_m.p, z.data(), nparam * sizeof(float), hipMemcpyHostToDevice));
  NED_HIP(hipMemcpy(adam_v.p, z.data(), nparam * sizeof(float), hipMemcpyHostToDevice));
  adam_t = 0;
}

void Mlp::ensure_rows(int rows) {
  if (rows <= cap_rows) return;
  h1.ensure((size_t)rows * w * sizeof(float));
  h2.ensure((size_t)rows * w * sizeof(float));
  yb.ensure((size_t)rows * n_out * sizeof(float));
  dyb.ensure((size_t)rows * n_out * sizeof(float));
  dh.ensure((size_t)rows * w * sizeof(float));
  part.ensure((size_t)(nb((long long)rows * n_out) + 1) * sizeof(double));
  lossd.ensure(sizeof(double));
  cap_rows = rows;
}

void Mlp::forward(const float *X, int rows, float *Y, hipStream_t s) {
  ensure_rows(rows);
  float *P = params.as<float>();
  float *H1 = h1.as<float>(), *H2 = h2.as<float>();
  mm_nt(X, P + off[0], H1, rows, w, n_in, s);
  hipLaunchKernelGGL(kern_bias_act, dim3(nb((long long)rows * w)), dim3(256), 0, s, H1, P + off[1], (long long)rows, w, 1);
  mm_nt(H1, P + off[2], H2, rows, w, w, s);
  hipLaunchKernelGGL(kern_bias_act, dim3(nb((long long)rows * w)), dim3(256), 0, s, H2, P + off[3], (long long)rows, w, 1);
  mm_nt(H2, P + off[4], Y, rows, n_out, w, s);
  hipLaunchKernelGGL(kern_bias_act, dim3(nb((long long)rows * n_out)), dim3(256), 0, s, Y, P + off[5], (long long)rows,
                     n_out, 0);
  check_launch("mlp forward");
}

void Mlp::forward_bf16(const float *X, int rows, float *Y, hipStream_t s) {
  require(n_in % 64 == 0 && w % 128 == 0 && n_out % 128 == 0,
          "mlp: BF16 inference needs n_in a multiple of 64, width and n_out of 128");
  const int rp = (rows + 127) / 128 * 128;
  const long long nw = (long long)w * n_in + (long long)w * w + (long long)n_out * w;
  if (rp > cap_pad) {
    abf.ensure((size_t)rp * std::max(n_in, w) * sizeof(uint16_t));
    pad_out.ensure((size_t)rp * std::max(w, n_out) * sizeof(float));
    cap_pad = rp;
  }
  const size_t nz = std::max<size_t>((size_t)nw, (size_t)cap_pad * std::max(n_in, w));
  if (zero.bytes < nz * sizeof(uint16_t)) {  // a fill kernel (rows = 0: every entry padding)
    zero.ensure(nz * sizeof(uint16_t));
    hipLaunchKernelGGL(kern_to_bf16, dim3(nb((long long)nz)), dim3(256), 0, s, nullptr, 0LL, (long long)nz, 1,
                       zero.as<uint16_t>());
  }
  float *P = params.as<float>();
  if (bf16_version != version) {  // the weights as bf16, once per parameter update
    wbf.ensure((size_t)nw * sizeof(uint16_t));
    uint16_t *W = wbf.as<uint16_t>();
    const long long n1 = (long long)w * n_in, n2 = (long long)w * w, n3 = (long long)n_out * w;
    hipLaunchKernelGGL(kern_to_bf16, dim3(nb(n1)), dim3(256), 0, s, P + off[0], (long long)w, (long long)w, n_in, W);
    hipLaunchKernelGGL(kern_to_bf16, dim3(nb(n2)), dim3(256), 0, s, P + off[2], (long long)w, (long long)w, w, W + n1);
    hipLaunchKernelGGL(kern_to_bf16, dim3(nb(n3)), dim3(256), 0, s, P + off[4], (long long)n_out, (long long)n_out, w,
                       W + n1 + n2);
    bf16_version = version;
  }
  const uint16_t *W1 = wbf.as<uint16_t>(), *W2 = W1 + (long long)w * n_in, *W3 = W2 + (long long)w * w;
  const uint16_t *Z = zero.as<uint16_t>();
  uint16_t *A = abf.as<uint16_t>();
  float *O = pad_out.as<float>();
  // O (rp x out) = A Wl^T: the split-bf16 product with zero low planes, which is
  // exactly bf16 x bf16 with FP32 accumulation
  auto mm = [&](int k, const uint16_t *Wl, int out) {
    gemm_nt_bf16x3(NtOp{A, Z, k, 0}, NtOp{Wl, Z, k, 0}, O, out, 0, rp, out, k, 1, 1.0f, 0, s, false);
  };
  hipLaunchKernelGGL(kern_to_bf16, dim3(nb((long long)rp * n_in)), dim3(256), 0, s, X, (long long)rows, (long long)rp,
                     n_in, A);
  mm(n_in, W1, w);
  hipLaunchKernelGGL(kern_bias_tanh_bf16, dim3(nb((long long)rp * w)), dim3(256), 0, s, O, P + off[1], (long long)rows,
                     (long long)rp, w, A);
  mm(w, W2, w);
  hipLaunchKernelGGL(kern_bias_tanh_bf16, dim3(nb((long long)rp * w)), dim3(256), 0, s, O, P + off[3], (long long)rows,
                     (long long)rp, w, A);
  mm(w, W3, n_out);
  hipLaunchKernelGGL(kern_bias_rows, dim3(nb((long long)rows * n_out)), dim3(256), 0, s, O, P + off[5], (long long)rows,
                     n_out, Y);
  check_launch("mlp forward (bf16)");
}

double Mlp::loss_grad(const float *X, const float *T, int rows, hipStream_t s) {
  ensure_rows(rows);
  float *P = params.as<float>(), *Gd = grads.as<float>();
  float *H1 = h1.as<float>(), *H2 = h2.as<float>(), *Y = yb.as<float>(), *dY = dyb.as<float>(), *dH = dh.as<float>();
  forward(X, rows, Y, s);
  const long long ny = (long long)rows * n_out;
  const unsigned nbl = nb(ny);
  hipLaunchKernelGGL(kern_mse, dim3(nbl), dim3(256), 0, s, Y, T, ny, 1.0f / (float)ny, dY, part.as<double>());
  hipLaunchKernelGGL(kern_sum_d, dim3(1), dim3(256), 0, s, part.as<double>(), (int)nbl, lossd.as<double>());
  // layer 3
  mm_tn(dY, H2, Gd + off[4], n_out, w, rows, s);
  hipLaunchKernelGGL(kern_colsum, dim3(nb(n_out)), dim3(256), 0, s, dY, (long long)rows, n_out, Gd + off[5]);
  mm_nn(dY, P + off[4], dH, rows, w, n_out, s);
  hipLaunchKernelGGL(kern_tanh_back, dim3(nb((long long)rows * w)), dim3(256), 0, s, dH, H2, (long long)rows * w);
  // layer 2 (dH now holds dZ2; dZ1 goes to H2's storage, which is no longer needed)
  mm_tn(dH, H1, Gd + off[2], w, w, rows, s);
  hipLaunchKernelGGL(kern_colsum, dim3(nb(w)), dim3(256), 0, s, dH, (long long)rows, w, Gd + off[3]);
  mm_nn(dH, P + off[2], H2, rows, w, w, s);
  hipLaunchKernelGGL(kern_tanh_back, dim3(nb((long long)rows * w)), dim3(256), 0, s, H2, H1, (long long)rows * w);
  // layer 1
  mm_tn(H2, X, Gd + off[0], w, n_in, rows, s);
  hipLaunchKernelGGL(kern_colsum, dim3(nb(w)), dim3(256), 0, s, H2, (long long)rows, w, Gd + off[1]);
  check_launch("mlp backward");
  double loss = 0;
  NED_HIP(hipMemcpyAsync(&loss, lossd.p, sizeof(double), hipMemcpyDeviceToHost, s));
  NED_HIP(hipStreamSynchronize(s));
  return loss / (double)ny;
}

void Mlp::adam(float lr, hipStream_t s) {
  adam_t++;
  version++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)adam_t), c2 = 1.0f - std::pow(b2, (float)adam_t);
  hipLaunchKernelGGL(kern_adam, dim3(nb(nparam)), dim3(256), 0, s, params.as<float>(), grads.as<float>(),
                     adam_m.as<float>(), adam_v.as<float>(), nparam, lr, b1, b2, 1e-8f, c1, c2);
  check_launch("mlp adam");
}

void Mlp::train(const float *X_host, const float *T_host, int rows, int epochs, int batch, float lr,
                unsigned long long seed, float *loss_out) {
  require(rows > 0 && epochs > 0 && batch > 0, "mlp train: rows, epochs and batch must be positive");
  batch = std::min(batch, rows);
  DevBuf dX, dT, dIdx, bX, bT;
  dX.ensure((size_t)rows * n_in * sizeof(float));
  dT.ensure((size_t)rows * n_out * sizeof(float));
  NED_HIP(hipMemcpyAsync(dX.p, X_host, (size_t)rows * n_in * sizeof(float), hipMemcpyHostToDevice, stream));
  NED_HIP(hipMemcpyAsync(dT.p, T_host, (size_t)rows * n_out * sizeof(float), hipMemcpyHostToDevice, stream));
  dIdx.ensure((size_t)rows * sizeof(int));
  bX.ensure((size_t)batch * n_in * sizeof(float));
  bT.ensure((size_t)batch * n_out * sizeof(float));
  std::mt19937_64 rng(seed);
  std::vector<int> idx(rows);
  for (int e = 0; e < epochs; e++) {
    for (int i = 0; i < rows; i++) idx[i] = i;
    std::shuffle(idx.begin(), idx.end(), rng);
    NED_HIP(hipMemcpyAsync(dIdx.p, idx.data(), rows * sizeof(int), hipMemcpyHostToDevice, stream));
    double sum = 0;
    int nbat = 0;
    for (int r0 = 0; r0 + batch <= rows; r0 += batch) {
      hipLaunchKernelGGL(kern_gather_rows, dim3(nb((long long)batch * n_in)), dim3(256), 0, stream, dX.as<float>(),
                         dIdx.as<int>() + r0, batch, n_in, bX.as<float>());
      hipLaunchKernelGGL(kern_gather_rows, dim3(nb((long long)batch * n_out)), dim3(256), 0, stream, dT.as<float>(),
                         dIdx.as<int>() + r0, batch, n_out, bT.as<float>());
      sum += loss_grad(bX.as<float>(), bT.as<float>(), batch, stream);
      adam(lr, stream);
      nbat++;
    }
    if (loss_out) loss_out[e] = (float)(sum / std::max(nbat, 1));
  }
  NED_HIP(hipStreamSynchronize(stream));
}

}  // namespace ned
