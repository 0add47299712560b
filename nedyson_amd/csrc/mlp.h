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

  Mlp(int in, int width, int out, int device, unsigned long long seed);
  ~Mlp();
  Mlp(const Mlp &) = delete;
  Mlp &operator=(const Mlp &) = delete;
  void reset_adam();
  void ensure_rows(int rows);
  // Y (rows x n_out) from X (rows x n_in), device pointers, on stream s
  void forward(const float *X, int rows, float *Y, hipStream_t s);
  // mean-squared loss of the rows against T, gradients into grads
  double loss_grad(const float *X, const float *T, int rows, hipStream_t s);
  void adam(float lr, hipStream_t s);
  // minibatch Adam over host data; loss_out[e] = mean batch loss of epoch e
  void train(const float *X_host, const float *T_host, int rows, int epochs, int batch, float lr,
             unsigned long long seed, float *loss_out);
};

}  // namespace ned
