"""The learned initial guess: GPU forward, backward and Adam kernels against
the CPU oracle, and what the trained network gains over the polynomial
predictor on GF2 trajectories the method itself generates."""
import numpy as np
import pytest

import nedyson_amd as ned
from nedyson_amd import predictor, problem
from oracle import mlp as M

pytestmark = pytest.mark.gpu


def _oracle_params(net):
    return [x.astype(np.float64) for x in net.split(net.get_params())]


def test_mlp_forward_backward_vs_oracle():
    rng = np.random.default_rng(1)
    net = ned.Mlp(48, 64, 16, seed=3)
    X = rng.normal(size=(100, 48)).astype(np.float32)
    T = rng.normal(size=(100, 16)).astype(np.float32)
    loss, g = net.loss_grad(X, T)
    P = _oracle_params(net)
    loss_ref, g_ref = M.loss_and_grads(P, X.astype(np.float64), T.astype(np.float64))
    assert abs(loss - loss_ref) < 1e-5 * loss_ref
    for a, b in zip(net.split(g), g_ref):
        assert np.max(np.abs(a - b)) < 1e-4 * np.max(np.abs(b)), np.max(np.abs(a - b))
    Y = net.predict(X)
    Y_ref, _ = M.forward(P, X.astype(np.float64))
    assert np.max(np.abs(Y - Y_ref)) < 1e-4 * np.max(np.abs(Y_ref))


def test_mlp_adam_step_vs_oracle():
    rng = np.random.default_rng(2)
    net = ned.Mlp(20, 32, 8, seed=5)
    X = rng.normal(size=(40, 20)).astype(np.float32)
    T = rng.normal(size=(40, 8)).astype(np.float32)
    P = _oracle_params(net)
    m = [np.zeros_like(x) for x in P]
    v = [np.zeros_like(x) for x in P]
    for step in (1, 2, 3):
        _, g = net.loss_grad(X, T)
        net.adam_step(1e-2)
        _, g_ref = M.loss_and_grads(P, X.astype(np.float64), T.astype(np.float64))
        M.adam_step(P, g_ref, m, v, step, 1e-2)
        for a, b in zip(net.split(net.get_params()), P):
            assert np.max(np.abs(a - b)) < 1e-4, (step, np.max(np.abs(a - b)))


def test_mlp_training_reduces_loss():
    """A smooth map from 768 inputs to 128 outputs: the minibatch trainer
    brings the loss down by orders of magnitude."""
    rng = np.random.default_rng(3)
    A = rng.normal(size=(768, 128)).astype(np.float32) / np.sqrt(768)
    X = rng.normal(size=(4096, 768)).astype(np.float32)
    Y = np.tanh(X @ A).astype(np.float32)
    net = ned.Mlp(seed=7)
    losses = net.train(X, Y, epochs=30, batch=256, lr=1e-3, seed=1)
    assert losses[-1] < 0.05 * losses[0], losses


def _system(seed, norm, nt):
    return problem.Problem(nao=4, naux=12, nelec=2, beta=10.0, ntau=15, nt=nt, dt=0.02, seed=seed, spacing=1.0,
                           b_sd=0.1, quench="random", quench_norm=norm)


def _context(p):
    c = ned.Context(p.nao, p.naux, p.nt, p.ntau, p.k)
    c.set_coulomb(p.B)
    c.upload(ned.NED_G, p.initial_green())
    c.set_matsubara(p.beta)
    c.do_boot(p.fock, p.mu, p.beta, p.dt, "f64", 50, 1e-12)
    return c


def _run(c, p, nt, tol):
    """do_tstp with one corrector pass, then plain passes until the update norm
    is below tol: per step (first-pass update norm, passes used)."""
    out = []
    for T in range(p.k + 1, nt + 1):
        _, err = c.do_tstp(T, p.h0, p.mu, p.beta, p.dt, "f64", 1, 0.0)
        first, passes = err, 1
        while err >= tol and passes < 30:
            c.hf(T, p.h0)
            c.sigma(T, "f64")
            err = float(np.sum(c.dyson_step(T, p.mu, p.beta, p.dt)))
            passes += 1
        assert err < tol
        out.append((T, first, passes))
    return out


def test_predictor_gain_on_held_out_system():
    """Trajectories of four seeded nao = 4 GF2 systems (random quenches of norm
    0.1..0.4) train the network; on a fifth system the learned guess must
    (a) predict the polynomial extrapolation's error well enough to cut the
    guess error at the predictor points, offline, and (b) cut the corrector's
    first update norm on the steps where it is used, online, with no
    fallbacks.  The corrector passes to reach 1e-8 are reported for both."""
    nt = 110
    residuals = []
    for seed, norm in ((11, 0.1), (12, 0.2), (13, 0.3), (14, 0.4)):
        p = _system(seed, norm, nt)
        with _context(p) as c:
            r = predictor.collect(c, p, p.k + 1, nt)
        residuals.append(r)
        print(f"[predictor] trajectory seed={seed}: {len(r)} steps", flush=True)
    Xs, Ys = zip(*[predictor.dataset(r)[:2] for r in residuals])
    X, Y = np.concatenate(Xs), np.concatenate(Ys)
    net, losses = predictor.train(X, Y, width=512, epochs=150, batch=256, lr=1e-3, seed=0)
    print(f"[predictor] {X.shape[0]} rows, loss {losses[0]:.3e} -> {losses[-1]:.3e}", flush=True)

    # (a) offline on the held-out trajectory
    p = _system(15, 0.25, nt)
    with _context(p) as c:
        r = predictor.collect(c, p, p.k + 1, nt)
    Xt, Yt, st = predictor.dataset(r)
    pred = net.predict(Xt)
    e_poly = np.sqrt(np.mean((Yt * st) ** 2))
    e_mlp = np.sqrt(np.mean(((Yt - pred) * st) ** 2))
    print(f"[predictor] held-out guess error at the points: polynomial {e_poly:.3e}, learned {e_mlp:.3e}",
          flush=True)
    assert e_mlp < 0.5 * e_poly

    # (b) online, same system, both guesses
    res = {}
    for name, use in (("poly", False), ("mlp", True)):
        with _context(p) as c:
            if use:
                c.set_predictor(net)
            res[name] = [x for x in _run(c, p, nt, 1e-8) if x[0] >= p.k + 39]
            if use:
                used, fb = c.predictor_stats()
                assert used > 0 and fb == 0, (used, fb)
    first = {k: float(np.mean([x[1] for x in v])) for k, v in res.items()}
    passes = {k: int(sum(x[2] for x in v)) for k, v in res.items()}
    print(f"[predictor] mean first update norm {first}; corrector passes to 1e-8 {passes}", flush=True)
    assert first["mlp"] < 0.5 * first["poly"]
