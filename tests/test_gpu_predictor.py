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


def _bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("rows", [16, 300, 2304])
def test_mlp_bf16_inference(rows):
    """BF16 inference against the same network evaluated in float64 with the
    weights and every layer's input rounded to bf16 as the GPU does (products
    then differ only by FP32 accumulation order, and an activation can land
    on the other side of a bf16 rounding boundary), and against the FP32
    forward pass within the bf16 rounding error.  rows = 16 and 300 exercise
    the padding of the row count to whole 128-row tiles."""
    rng = np.random.default_rng(4)
    net = ned.Mlp(seed=11)
    X = rng.normal(size=(rows, 768)).astype(np.float32)
    Y32 = net.predict(X)
    net.set_bf16(True)
    Y16 = net.predict(X)
    W1, b1, W2, b2, W3, b3 = _oracle_params(net)
    h = np.tanh(_bf16(X) @ _bf16(W1).T + b1)
    h = np.tanh(_bf16(h) @ _bf16(W2).T + b2)
    Yref = _bf16(h) @ _bf16(W3).T + b3
    scale = np.max(np.abs(Yref))
    err_emul = np.max(np.abs(Y16 - Yref))
    err_32 = np.max(np.abs(Y16 - Y32))
    print(f"[mlp bf16] rows {rows}: max |y| {scale:.3f}, vs bf16 emulation {err_emul:.2e}, vs FP32 {err_32:.2e}",
          flush=True)
    assert err_emul < 2e-3 * scale
    assert np.mean(np.abs(Y16 - Yref)) < 1e-4 * scale
    assert err_32 < 5e-2 * scale
    # a parameter update refreshes the bf16 weights: W3 -> 2 W3 (exact in
    # bf16) gives y -> 2 y - b3
    P = net.get_params()
    P[net.split(np.arange(net.nparams))[4]] *= 2.0
    net.set_params(P)
    assert np.max(np.abs(net.predict(X) - (2 * Y16 - b3))) < 1e-5 * scale
    net.set_bf16(False)
    assert np.max(np.abs(net.predict(X) - (2 * Y32 - b3))) < 1e-5 * scale


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


def _gf2_system(i, nt):
    """System i of the learned-predictor configuration (nao = 48) with B
    entries of spread 0.05, from its GF2 equilibrium, bootstrapped in FP64."""
    p = problem.predictor_system(i, nt=nt, b_sd=0.05)
    with ned.Context(p.nao, p.naux, p.k, p.ntau, p.k) as c:
        c.set_coulomb(p.B)
        p.mu, p.fock, _, _ = c.matsubara_gf2(p.h1, p.beta, nelec=float(p.nelec), mu=p.mu, tol=1e-9)
        p.Gmat = c.read_blocks(ned.NED_G, ned.NED_MAT, 0, p.ntau + 1)
    return p


def _booted(p):
    c = ned.Context(p.nao, p.naux, p.nt, p.ntau, p.k)
    c.set_coulomb(p.B)
    c.write_blocks(ned.NED_G, ned.NED_MAT, 0, p.Gmat)
    c.set_matsubara(p.beta)
    c.g0_from_h0(p.k, p.fock, p.mu, p.beta, p.dt)
    c.do_boot(p.fock, p.mu, p.beta, p.dt, "f64", 60, 1e-10)
    return c


@pytest.mark.timeout(900)
def test_predictor_gain_nao48():
    """The learned-predictor configuration at reduced counts: four nao = 48
    training systems (GF2 equilibrium, random quench, FP64 corrector to 1e-10,
    80 steps) train the network on the GPU; on a held-out system, run three
    times from the same equilibrium with the corrector iterated to 1e-8, the
    learned guess (BF16 inference, and FP32) needs fewer corrector passes than
    the polynomial one over the steps where it acts, with no fallbacks.
    Passes and wall time are reported for each."""
    nt = 80
    X, Y = [], []
    for i in range(4):
        p = _gf2_system(i, nt)
        with _booted(p) as c:
            r = predictor.collect(c, p, p.k + 1, nt, corr=20, corr_tol=1e-10)
        x, y, _ = predictor.dataset(r)
        X.append(x)
        Y.append(y)
        print(f"[predictor48] training system {i}: quench norm {p.quench_norm:.3f}, {len(r)} steps", flush=True)
    X, Y = predictor.subsample(np.concatenate(X), np.concatenate(Y), 150000)
    net, losses = predictor.train(X, Y, width=512, epochs=40, batch=256, lr=1e-3, seed=0)
    print(f"[predictor48] {X.shape[0]} rows, loss {losses[0]:.3e} -> {losses[-1]:.3e}", flush=True)
    assert losses[-1] < 0.3 * losses[0]

    p = _gf2_system(problem.PREDICTOR_NSYS - 1, nt)
    t1 = predictor.first_learned_step(p.k)
    res = {}
    for mode, bf16 in (("polynomial", None), ("learned_bf16", True), ("learned_fp32", False)):
        with _booted(p) as c:
            if bf16 is not None:
                net.set_bf16(bf16)
                c.set_predictor(net)
            steps = [s for s in predictor.run_to_tolerance(c, p, p.k + 1, nt) if s[0] >= t1]
            stats = c.predictor_stats() if bf16 is not None else (0, 0)
            c.set_predictor(None)
        res[mode] = (sum(s[1] for s in steps), sum(s[2] for s in steps), stats)
        print(f"[predictor48] {mode}: {res[mode][0]} passes over {len(steps)} steps, {res[mode][1]:.2f} s, "
              f"used/fallbacks {stats}", flush=True)
    for mode in ("learned_bf16", "learned_fp32"):
        used, fb = res[mode][2]
        assert used >= nt - t1 + 1 and fb == 0, (mode, used, fb)
        assert res[mode][0] <= 0.95 * res["polynomial"][0], res
