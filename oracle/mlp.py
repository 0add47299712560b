"""CPU oracle of the next-slice predictor network.

TEST INFRASTRUCTURE ONLY: the checker for the GPU forward, backward and Adam
kernels of the predictor (the product never imports it).  The network has no
counterpart in the reference -- it replaces the polynomial predictor
dyson::Extrapolate (src/dyson.cpp:47-117) as the corrector's initial guess --
so it is pinned by its own mathematics: central finite differences of the
loss check the backward pass here, and the GPU is checked against this file.

Network: x (in) -> tanh(W1 x + b1) -> tanh(W2 h1 + b2) -> W3 h2 + b3 (out),
rows of a batch independent; loss = mean over every output entry of
(y - t)^2.  Weights are row-major (out_features, in_features).
"""
import numpy as np


def forward(params, X):
    W1, b1, W2, b2, W3, b3 = params
    Z1 = X @ W1.T + b1
    H1 = np.tanh(Z1)
    Z2 = H1 @ W2.T + b2
    H2 = np.tanh(Z2)
    Y = H2 @ W3.T + b3
    return Y, (X, H1, H2)


def loss_and_grads(params, X, T):
    """Mean-squared loss and its gradient with respect to every parameter."""
    W1, b1, W2, b2, W3, b3 = params
    Y, (X, H1, H2) = forward(params, X)
    n = Y.size
    loss = float(np.mean((Y - T) ** 2))
    dY = 2.0 * (Y - T) / n
    dW3 = dY.T @ H2
    db3 = dY.sum(axis=0)
    dH2 = dY @ W3
    dZ2 = dH2 * (1.0 - H2 ** 2)
    dW2 = dZ2.T @ H1
    db2 = dZ2.sum(axis=0)
    dH1 = dZ2 @ W2
    dZ1 = dH1 * (1.0 - H1 ** 2)
    dW1 = dZ1.T @ X
    db1 = dZ1.sum(axis=0)
    return loss, [dW1, db1, dW2, db2, dW3, db3]


def adam_step(params, grads, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1), in place."""
    for p, g, mm, vv in zip(params, grads, m, v):
        mm *= beta1
        mm += (1 - beta1) * g
        vv *= beta2
        vv += (1 - beta2) * g * g
        mhat = mm / (1 - beta1 ** step)
        vhat = vv / (1 - beta2 ** step)
        p -= lr * mhat / (np.sqrt(vhat) + eps)


def numerical_grad(params, X, T, which, idx, h=1e-6):
    """Central difference of the loss in one parameter entry (float64)."""
    p = params[which]
    old = p[idx]
    p[idx] = old + h
    lp, _ = loss_and_grads(params, X, T)
    p[idx] = old - h
    lm, _ = loss_and_grads(params, X, T)
    p[idx] = old
    return (lp - lm) / (2 * h)
