"""Training data and training loop for the learned initial guess.

The network (nedyson_amd.Mlp, hand-written GPU kernels) maps, per orbital
pair, the extrapolation errors of the last six steps at the 2 x 32 predictor
points (retarded row and lesser column; 16 columns after the quench and the 16
next to the diagonal; real and imaginary parts) to the error of the current
step.  Each (component, zone) group of inputs and targets is scaled by the RMS
of that group's inputs, as the GPU feature kernel does at inference time.

Trajectories come from the method itself: run ned_do_tstp with the
polynomial predictor and read each step's errors with
Context.predictor_residual(T).
"""
import time

import numpy as np

from . import PRED_INPUTS, PRED_OUTPUTS, Mlp

HIST = 6
GROUPS = 4  # (component, zone) groups of 32 features, each with its own scale


def first_learned_step(k):
    """The first step whose guess can use the network: the extrapolation
    errors are recorded from step k + 33 on (the 16 quench columns and the 16
    next to the diagonal must not overlap) and the network needs six."""
    return k + 33 + HIST


def dataset(residuals):
    """residuals: list of (n2, 128) float32 arrays of consecutive steps.
    Returns X (rows, 768), Y (rows, 128) and the scale of every output
    (rows, 128), one row per orbital pair and predicted step; each
    (component, zone) group of 32 features is divided by its RMS over the six
    input steps."""
    R = [np.asarray(r, dtype=np.float32) for r in residuals]
    Xs, Ys, Ss = [], [], []
    for t in range(HIST, len(R)):
        h = np.stack([R[t - 1 - q] for q in range(HIST)], axis=1)  # (n2, 6, 128), newest first
        g = h.reshape(h.shape[0], HIST, GROUPS, -1).astype(np.float64)
        sig = np.sqrt(np.mean(g ** 2, axis=(1, 3))).astype(np.float32) + np.float32(1e-30)  # (n2, 4)
        scale = np.repeat(sig, PRED_OUTPUTS // GROUPS, axis=1)  # (n2, 128)
        Xs.append((h / scale[:, None, :]).reshape(h.shape[0], -1))
        Ys.append(R[t] / scale)
        Ss.append(scale)
    if not Xs:
        return (np.zeros((0, PRED_INPUTS), np.float32), np.zeros((0, PRED_OUTPUTS), np.float32),
                np.zeros((0, PRED_OUTPUTS), np.float32))
    return np.concatenate(Xs), np.concatenate(Ys), np.concatenate(Ss)


def collect(ctx, p, t_from, t_to, prec="f64", corr=3, corr_tol=0.0):
    """Run do_tstp for t_from..t_to on a prepared context and return the list
    of per-step extrapolation errors (steps with T >= k + 33 only)."""
    out = []
    for T in range(t_from, t_to + 1):
        ctx.do_tstp(T, p.h0, p.mu, p.beta, p.dt, prec, corr, corr_tol)
        if T >= p.k + 33:
            out.append(ctx.predictor_residual(T))
    return out


def subsample(X, Y, rows, seed=0):
    """At most `rows` rows of (X, Y), drawn without replacement (seeded)."""
    if X.shape[0] <= rows:
        return X, Y
    sel = np.sort(np.random.default_rng(seed).choice(X.shape[0], rows, replace=False))
    return X[sel], Y[sel]


def run_to_tolerance(ctx, p, t_from, t_to, prec="f64", corr_tol=1e-8, corr_max=20):
    """do_tstp for t_from..t_to with the corrector iterated until its update
    norm is below corr_tol (at most corr_max passes).  Returns per step
    (T, passes, seconds); the wall time includes the guess (polynomial, and
    the network where the context has one) and every pass."""
    out = []
    for T in range(t_from, t_to + 1):
        t0 = time.perf_counter()
        it, err = ctx.do_tstp(T, p.h0, p.mu, p.beta, p.dt, prec, corr_max, corr_tol)
        out.append((T, it, time.perf_counter() - t0))
        if not err < corr_tol:
            raise RuntimeError(f"step {T}: corrector update {err:.3e} after {it} passes")
    return out


def train(X, Y, width=512, epochs=200, batch=256, lr=1e-3, seed=0, device=0):
    net = Mlp(PRED_INPUTS, width, PRED_OUTPUTS, device=device, seed=seed)
    losses = net.train(X, Y, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return net, losses
