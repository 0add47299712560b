// The next-slice predictor network (see mlp.hip).
#pragma once
#include "ctx.h"

namespace ned {

struct Mlp {
  int n_in, w, n_out, device;
  hipStream_t stream = nullptr;
  long long off[7];  // W1, b1, W2, b2, W3, b3 offsets in params; off[6] = count
  long long nparam;
  DevBuf params, grads, adam_m, adam_v;
  DevBuf h1, h2, yb, dyb, dh, part, lossd;
  int cap_rows = 0;
  long long adam_t = 0;
  // BF16 inference: the weights rounded to bf16 (W1 | W2 | W3, refreshed when
  // the parameters change), the activations of a batch in bf16 (rows padded to
  // 128), a zero plane, and the padded FP32 layer outputs
  bool bf16_infer = false;
  long long version = 0, bf16_version = -1;
  DevBuf wbf, abf, zero, pad_out;
  int cap_pad = 0;

  Mlp(int in, int width, int out, int device, unsigned long long seed);
  ~Mlp();
  Mlp(const Mlp &) = delete;
  Mlp &operator=(const Mlp &) = delete;
  void reset_adam();
  void ensure_rows(int rows);
  // Y (rows x n_out) from X (rows x n_in), device pointers, on stream s, in
  // FP32 (the training forward pass)
  void forward(const float *X, int rows, float *Y, hipStream_t s);
  // the same with the weights and every layer's input rounded to bf16 and the
  // products on the bf16 matrix cores with FP32 accumulation (bias, tanh and
  // outputs in FP32); needs n_in % 64 == 0 and width, n_out % 128 == 0
  void forward_bf16(const float *X, int rows, float *Y, hipStream_t s);
  // inference in the mode set by bf16_infer
  void infer(const float *X, int rows, float *Y, hipStream_t s) {
    if (bf16_infer) forward_bf16(X, rows, Y, s);
    else forward(X, rows, Y, s);
  }
  // mean-squared loss of the rows against T, gradients into grads
  double loss_grad(const float *X, const float *T, int rows, hipStream_t s);
  void adam(float lr, hipStream_t s);
  // minibatch Adam over host data; loss_out[e] = mean batch loss of epoch e
  void train(const float *X_host, const float *T_host, int rows, int epochs, int batch, float lr,
             unsigned long long seed, float *loss_out);
};

}  // namespace ned
