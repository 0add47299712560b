"""Pinning the CPU oracle (the checker of the GPU path) against the reference's
own golden data, brute-force definitions and closed forms."""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

from contour_fixtures import set_up_G, set_up_G_herm
from helpers import free_green, gf2_functional, poly_green, random_green
from nedyson_amd.green import GreenFunc

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _exact_inverse_vandermonde(k):
    n = k + 1
    A = [[Fraction(i) ** a for a in range(n)] + [Fraction(int(i == j)) for j in range(n)] for i in range(n)]
    for c in range(n):
        p = next(r for r in range(c, n) if A[r][c] != 0)
        A[c], A[p] = A[p], A[c]
        piv = A[c][c]
        A[c] = [v / piv for v in A[c]]
        for r in range(n):
            if r != c and A[r][c] != 0:
                f = A[r][c]
                A[r] = [x - f * y for x, y in zip(A[r], A[c])]
    return np.array([[float(A[i][n + j]) for j in range(n)] for i in range(n)])


def test_integrator_against_calccoeff(orc):
    """tests/integration-test.cpp:5-69 (golden tables of calccoeff.h5)."""
    cc = json.load(open(os.path.join(GOLD, "calccoeff.json")))
    I = orc.integrator(5)
    # PInterp: the golden file carries double-rounding noise of 1.8e-13 relative
    # to the exact rational inverse Vandermonde; the oracle is exact.
    exact = _exact_inverse_vandermonde(5).reshape(-1)
    assert np.max(np.abs(I["poly_interp"] - exact)) < 1e-15
    assert np.linalg.norm(np.array(cc["PInterp"]) - I["poly_interp"]) < 1e-12
    assert np.linalg.norm(np.array(cc["PDiff"]) - I["poly_diff"]) < 1e-12
    assert np.linalg.norm(np.array(cc["PInteg"]) - I["poly_integ"]) < 1e-12
    assert np.linalg.norm(np.array(cc["BD"]) - I["bd"]) < 1e-12
    W = orc.gregory_table(5, 11)
    assert np.linalg.norm(np.array(cc["Start"]) - W[:6, :6].reshape(-1)) < 1e-12
    rows = [W[i, j] for i in range(6, 12) for j in range(i + 1)]
    assert np.linalg.norm(np.array(cc["Sigma"]) - np.array(rows)) < 1e-12
    assert np.linalg.norm(np.array(cc["w"]) - I["Omega"][30:36]) < 1e-12


def test_integrator_closed_forms(orc):
    """Gregory weights integrate polynomials of degree <= k exactly; BD weights
    differentiate them exactly; extrapolation weights reproduce x=-1."""
    k = 5
    I = orc.integrator(k)
    W = orc.gregory_table(k, 40)
    for n in range(0, 41):
        top = max(n, k)
        for a in range(k + 1):
            s = sum(W[n, j] * float(j) ** a for j in range(top + 1))
            assert abs(s - n ** (a + 1) / (a + 1)) < 1e-9 * max(1.0, n ** (a + 1))
    bd = I["bd"]
    for a in range(k + 2):
        # d/dt t^a at t = T from samples T, T-1, ... (T = 10)
        T = 10.0
        s = sum(bd[l] * (T - l) ** a for l in range(k + 2))
        assert abs(s - (a * T ** (a - 1) if a else 0.0)) < 1e-8 * max(1, T ** a)
    ex = I["ex"]
    for a in range(k + 1):
        assert abs(sum(ex[j] * float(j) ** a for j in range(k + 1)) - (-1.0) ** a) < 1e-10


def test_gll_nodes_and_mass(orc):
    """GLL nodes are the roots of (1-x^2) P_N'(x); the mass matrix integrates
    products of degree-N polynomials exactly."""
    from numpy.polynomial import legendre as L
    for N in (5, 30, 63):
        x, mass = orc.gll(N)
        assert x[0] == -1 and x[-1] == 1 and np.all(np.diff(x) > 0)
        dP = L.legder([0] * N + [1])
        assert np.max(np.abs(L.legval(x[1:-1], dP))) < 1e-9 * N * N
        assert np.max(np.abs(x + x[::-1])) < 1e-14
        for a in range(0, N + 1, max(1, N // 4)):
            for b in range(0, N + 1, max(1, N // 3)):
                va, vb = x ** a, x ** b
                exact = 0.0 if (a + b) % 2 else 2.0 / (a + b + 1)
                assert abs(va @ mass @ vb - exact) < 1e-12


def test_integrals_golden(orc):
    """tests/integrals-test.cpp:8-212 golden scalars."""
    gd = json.load(open(os.path.join(GOLD, "integrals.json")))
    P = gd["params"]
    Nt, Ntau, beta, dt = P["Nt"], P["Ntau"], P["beta"], P["dt"]
    x, _ = orc.gll(Ntau)
    D = orc.Dyson(Nt, Ntau, 2, 5)
    G, S = set_up_G_herm(x, Nt, Ntau, beta, dt)
    for t, re, im, tol in gd["energy_conv"]:
        assert abs(D.energy_conv(t, S, G, beta, dt) - re) < tol, t
    G, Gcc, S, Scc = set_up_G(x, Nt, Ntau, beta, dt)
    for n, re, im, tol in gd["CTV1"]:
        assert abs(D.ctv1(S, Scc, G, n, dt).sum() - (re + 1j * im)) < tol, n
    for m, re, im, tol in gd["Cles2_tstp"]:
        assert abs(D.cles2(G, Gcc, S, Scc, m, dt)[:max(m, 5) + 1].sum() - (re + 1j * im)) < tol, m
    for m, re, im, tol in gd["Cles3_tstp"]:
        assert abs(D.cles3(S, Scc, G, Gcc, m, beta)[:max(m, 5) + 1].sum() - (re + 1j * im)) < tol, m
    for n, re, im, tol in gd["Ctv_tstp"]:
        G, Gcc, S, Scc = set_up_G(x, Nt, Ntau, beta, dt)
        D.ctv_tstp(n, G, G, Gcc, S, Scc, beta, dt)
        assert abs(G.tv[n].sum() - (re + 1j * im)) < tol, n


def test_mixing_kernel_closed_form(orc):
    """Conv.mixing on polynomials: int_0^beta A(t) B(t - tau) dt with the
    antiperiodic continuation, against adaptive quadrature of the same
    polynomials."""
    from scipy.integrate import quad
    N, beta, sig = 12, 2.5, -1
    x, _ = orc.gll(N)
    taus = beta * (x + 1) / 2
    a_c = [0.3, -1.0, 0.5, 0.2]
    b_c = [1.0, 0.1, -0.4, 0.05, 0.01]
    A = np.polyval(a_c, taus).reshape(-1, 1, 1).astype(complex)
    Bv = np.polyval(b_c, taus).reshape(-1, 1, 1).astype(complex)
    K = orc.mixing_kernel(N, Bv, beta, sig)
    C = np.einsum("j,ij->i", A[:, 0, 0], K[:, :, 0, 0])
    for i, ti in enumerate(taus):
        f1 = quad(lambda t: np.polyval(a_c, t) * np.polyval(b_c, t - ti), ti, beta)[0]
        f2 = quad(lambda t: np.polyval(a_c, t) * np.polyval(b_c, t - ti + beta), 0, ti)[0]
        assert abs(C[i] - (f1 + sig * f2)) < 1e-12


def test_sigma_gemm_chain_vs_brute_force(orc):
    """tests/mol2bSolver-test.cpp:80-132: solve == solve_loop (random V and G)."""
    rng = np.random.default_rng(7)
    nao, nalpha, nt, ntau = 2, 3, 4, 5
    V = rng.uniform(-1, 1, (nao, nao, nalpha))
    G = random_green(rng, nt, ntau, nao)
    S, SL = GreenFunc(nt, ntau, nao), GreenFunc(nt, ntau, nao)
    for t in range(nt + 1):
        orc.sigma_solve(V, t, S, G)
        orc.sigma_solve_loop(V, t, SL, G)
    for c in ("ret", "les", "tv"):
        assert np.linalg.norm(getattr(S, c) - getattr(SL, c)) < 1e-10


def test_sigma_against_definition(orc):
    """The restated GEMM chain against the second-Born functional written with
    the full U tensor (the cross-check of tests/mol2bSolver-test.cpp:229-237),
    for every component: Sigma^< = F[G^<, G^>], Sigma^R = F[G^>,G^<] - F[G^<,G^>]
    (Langreth), Sigma^rceil = F[G^rceil, G^lceil]."""
    rng = np.random.default_rng(11)
    nao, na, nt, ntau = 3, 4, 3, 4
    V = rng.uniform(-1, 1, (nao, nao, na))
    G = random_green(rng, nt, ntau, nao)
    S = GreenFunc(nt, ntau, nao)
    T = nt
    orc.sigma_solve(V, T, S, G)
    for t in range(T + 1):
        L, R = G.lesptr(t, T), G.retptr(T, t)
        Gt = R - L.conj().T
        assert np.max(np.abs(S.lesptr(t, T) - gf2_functional(V, L, Gt))) < 1e-11
        sr = gf2_functional(V, Gt, L) - gf2_functional(V, -L.conj().T, L - R.conj().T)
        assert np.max(np.abs(S.retptr(T, t) - sr)) < 1e-11
    for m in range(ntau + 1):
        A = G.tv[T, m]
        Bm = -G.sig * G.tv[T, ntau - m].conj().T
        assert np.max(np.abs(S.tv[T, m] - gf2_functional(V, A, Bm))) < 1e-11


def test_hf_vs_loop(orc):
    """tests/molHFSolver-test.cpp:37-65."""
    rng = np.random.default_rng(3)
    nt, nao, na = 3, 2, 5
    V = rng.uniform(-1, 1, (nao, nao, na))
    rho = rng.uniform(-1, 1, (nao, nao)) + 1j * rng.uniform(-1, 1, (nao, nao))
    h = rng.uniform(-1, 1, (nt + 1, nao, nao)) + 1j * rng.uniform(-1, 1, (nt + 1, nao, nao))
    hl = h.copy()
    for t in range(nt + 1):
        orc.hf(V, t, h, rho)
        orc.hf(V, t, hl, rho, loop=True)
    assert np.linalg.norm(h - hl) < 1e-10


def test_extrapolate_exact_on_polynomials(orc):
    """dyson::Extrapolate (src/dyson.cpp:47-117) is exact for degree-k data."""
    rng = np.random.default_rng(5)
    nt, ntau, nao, k, dt = 20, 6, 2, 5, 0.05
    g = poly_green(nt, ntau, nao, dt, k, rng)
    D = orc.Dyson(nt, ntau, nao, k)
    for n in (k + 1, k + 3, 14, 20):
        h = g.copy()
        h.ret_row(n)[:] = 0
        h.les_col(n)[:] = 0
        h.tv[n] = 0
        D.extrapolate(n, h)
        for a, b in ((h.ret_row(n), g.ret_row(n)), (h.les_col(n), g.les_col(n)), (h.tv[n], g.tv[n])):
            assert np.max(np.abs(a - b)) < 1e-9 * max(1, np.max(np.abs(b)))


def _free_setup(nao=3, nt=30, ntau=24, beta=4.0, dt=0.02, seed=2):
    rng = np.random.default_rng(seed)
    h = rng.normal(0, 0.5, (nao, nao))
    h = h + h.T
    mu = 0.1
    return h, mu, beta, dt, nt, ntau


def _free_run_error(orc, dt, nt):
    h, mu, beta, _, _, ntau = _free_setup()
    nao, k = h.shape[0], 5
    x, _ = orc.gll(ntau)
    ex = free_green(h, mu, beta, dt, nt, x)
    G = ex.copy()
    for n in range(k + 1, nt + 1):
        G.ret_row(n)[:] = 0
        G.les_col(n)[:] = 0
        G.tv[n] = 0
    S = GreenFunc(nt, ntau, nao)
    hmf = np.repeat(h[None].astype(complex), nt + 1, axis=0)
    D = orc.Dyson(nt, ntau, nao, k)
    for n in range(k + 1, nt + 1):
        D.extrapolate(n, G)
        D.dyson_step(n, G, S, hmf, mu, beta, dt)
    return [np.max(np.abs(getattr(G, c) - getattr(ex, c))) for c in ("ret", "les", "tv")]


def test_dyson_step_free_limit(orc):
    """With Sigma = 0 and constant hmf the KBE are solved by the closed-form
    free propagator (src/dyson_free.h:222-297); dyson_step must reproduce it
    with the sixth-order convergence of the k=5 BDF/Gregory scheme."""
    e1 = _free_run_error(orc, 0.02, 30)
    e2 = _free_run_error(orc, 0.01, 60)
    for a, b in zip(e1, e2):
        assert b < 1e-9
        assert a / max(b, 1e-300) > 30  # ~2^6


def test_bootstrap_free_limit(orc):
    """dyson_start with Sigma = 0 from the Matsubara boundary condition gives
    the free G for t <= k (src/dyson_boot.h:93-221)."""
    h, mu, beta, dt, nt, ntau = _free_setup(nt=8)
    nao, k = h.shape[0], 5
    x, _ = orc.gll(ntau)
    ex = free_green(h, mu, beta, dt, nt, x)
    G = GreenFunc(nt, ntau, nao)
    G.mat[:] = ex.mat
    S = GreenFunc(nt, ntau, nao)
    hmf = np.repeat(h[None].astype(complex), nt + 1, axis=0)
    D = orc.Dyson(nt, ntau, nao, k)
    for _ in range(3):
        D.dyson_start(G, S, hmf, mu, beta, dt)
    for t in range(k + 1):
        assert np.max(np.abs(G.tv[t] - ex.tv[t])) < 1e-8
        assert np.max(np.abs(G.ret_row(t) - ex.ret_row(t))) < 1e-8
        assert np.max(np.abs(G.les_col(t) - ex.les_col(t))) < 1e-8


def test_g0_from_h0_matches_closed_form(orc):
    """G0_from_h0 (src/dyson_free.h:226-297) against the numpy closed form."""
    h, mu, beta, dt, nt, ntau = _free_setup(nt=8)
    nao, k = h.shape[0], 5
    x, _ = orc.gll(ntau)
    ex = free_green(h, mu, beta, dt, nt, x)
    G = GreenFunc(nt, ntau, nao)
    D = orc.Dyson(nt, ntau, nao, k)
    D.g0_from_h0(G, mu, h, beta, dt)
    D.g0_matsubara(G, mu, h, beta)
    assert np.max(np.abs(G.mat - ex.mat)) < 1e-12
    for t in range(k + 1):
        assert np.max(np.abs(G.tv[t] - ex.tv[t])) < 1e-12
        assert np.max(np.abs(G.ret_row(t) - ex.ret_row(t))) < 1e-12
        assert np.max(np.abs(G.les_col(t) - ex.les_col(t))) < 1e-12


@pytest.mark.parametrize("nthreads", [1, 4])
def test_sigma_threaded_bit_identical(orc, nthreads):
    rng = np.random.default_rng(9)
    nao, na, nt, ntau = 3, 5, 4, 6
    V = rng.uniform(-1, 1, (nao, nao, na))
    G = random_green(rng, nt, ntau, nao)
    S1, S2 = GreenFunc(nt, ntau, nao), GreenFunc(nt, ntau, nao)
    orc.sigma_solve(V, nt, S1, G, 1)
    orc.sigma_solve(V, nt, S2, G, nthreads)
    for c in ("ret", "les", "tv"):
        assert np.array_equal(getattr(S1, c), getattr(S2, c))


def test_mlp_oracle_backward_matches_finite_differences():
    """The predictor network's CPU oracle: analytic gradients (backward pass)
    against central differences of the loss, in float64."""
    from oracle import mlp as M
    rng = np.random.default_rng(3)
    n_in, w, n_out, rows = 7, 9, 4, 5
    params = [rng.normal(0, 0.5, (w, n_in)), rng.normal(0, 0.1, w), rng.normal(0, 0.5, (w, w)),
              rng.normal(0, 0.1, w), rng.normal(0, 0.5, (n_out, w)), rng.normal(0, 0.1, n_out)]
    X = rng.normal(size=(rows, n_in))
    T = rng.normal(size=(rows, n_out))
    _, grads = M.loss_and_grads(params, X, T)
    for which in range(6):
        for _ in range(4):
            idx = tuple(rng.integers(0, s) for s in params[which].shape)
            num = M.numerical_grad(params, X, T, which, idx)
            assert abs(num - grads[which][idx]) < 1e-7 * max(1.0, abs(num)), (which, idx)


def test_mlp_oracle_adam_matches_closed_form_first_step():
    """First Adam step moves every parameter by lr * sign(g) (bias-corrected
    moments m = g, v = g^2), up to eps."""
    from oracle import mlp as M
    rng = np.random.default_rng(4)
    p = [rng.normal(size=(3, 2))]
    g = [rng.normal(size=(3, 2))]
    m, v = [np.zeros((3, 2))], [np.zeros((3, 2))]
    p0 = p[0].copy()
    M.adam_step(p, g, m, v, 1, 0.01)
    assert np.allclose(p[0], p0 - 0.01 * np.sign(g[0]), atol=1e-8)


def test_dyson_step_ret_against_reference_timesteps(orc):
    """The retarded rows of the reference's stored GF propagation
    (tests/data/dyson_test_data.h5, Timesteps with Sigma and a time-dependent
    mean field) for T = k+1..20, each from the stored history: 1e-12."""
    import reference_fixtures as RF
    S, Gt, h = RF.green("Sigma"), RF.green("Timesteps"), RF.hmf()
    D = orc.Dyson(RF.NT, RF.NTAU, RF.NAO, RF.K)
    for T in range(RF.K + 1, RF.NT + 1):
        G = RF.green("Timesteps")
        D.dyson_step(T, G, S, h, RF.MU, RF.BETA, RF.DT)
        err = max(np.abs(G.retptr(T, j) - Gt.retptr(T, j)).max() for j in range(T + 1))
        assert err < 1e-12, (T, err)


def test_extrapolate_against_reference(orc):
    """dyson::Extrapolate against the reference's stored extrapolation run
    (dyson_test_data.h5, Extrap): every slice T = k+1..20 from the slices
    before it, all three components, 1e-12 relative."""
    import reference_fixtures as RF
    Ge = RF.green("Extrap")
    D = orc.Dyson(RF.NT, RF.NTAU, RF.NAO, RF.K)
    for T in range(RF.K + 1, RF.NT + 1):
        G = RF.green("Extrap")
        D.extrapolate(T, G)
        for a, b in ((G.ret_row(T), Ge.ret_row(T)), (G.les_col(T), Ge.les_col(T)), (G.tv[T], Ge.tv[T])):
            assert np.max(np.abs(a - b)) <= 1e-12 * max(1.0, np.max(np.abs(b))), T


def test_g0_from_h0_against_reference_free_test(orc):
    """G0_from_h0 against tests/data/Gfree_const.h5, the golden file of
    tests/free-test.cpp:8-44: G^R and G^< of steps 0..k directly; G^M and
    G^tv through the closed form, which the file holds on a uniform tau grid."""
    import reference_fixtures as RF
    ref = RF.green("Gfree")
    G = GreenFunc(RF.NT, RF.NTAU, RF.NAO)
    D = orc.Dyson(RF.NT, RF.NTAU, RF.NAO, RF.K)
    D.g0_from_h0(G, RF.MU, RF.H0_FREE, RF.BETA, RF.DT)
    nb = (RF.K + 1) * (RF.K + 2) // 2
    assert np.max(np.abs(G.ret[:nb] - ref.ret[:nb])) < 1e-12
    assert np.max(np.abs(G.les[:nb] - ref.les[:nb])) < 1e-12
    # the closed form on the fixture's uniform grid reproduces its G^M and G^tv
    from nedyson_amd.problem import fermi
    lam, V = np.linalg.eigh(RF.MU * np.eye(RF.NAO) - RF.H0_FREE)
    taus = np.linspace(0, RF.BETA, RF.NTAU + 1)

    def fexp(t, w):
        return np.where(w < 0, np.exp(w * t) * fermi(RF.BETA, w), np.exp((t - RF.BETA) * w) * fermi(RF.BETA, -w))
    gm = np.array([-(V * fexp(t, lam)) @ V.T for t in taus])
    assert np.max(np.abs(gm - ref.mat)) < 1e-12
    for m in range(RF.K + 1):
        U = V @ np.diag(np.exp(1j * RF.DT * m * lam)) @ V.T
        tv = np.array([1j * U @ ((V * fexp(t, -lam)) @ V.T) for t in taus])
        assert np.max(np.abs(tv - ref.tv[m])) < 1e-12
    # and the oracle's G^tv on its own (GLL) nodes is the same closed form
    x, _ = orc.gll(RF.NTAU)
    for m in range(RF.K + 1):
        U = V @ np.diag(np.exp(1j * RF.DT * m * lam)) @ V.T
        tv = np.array([1j * U @ ((V * fexp(t, -lam)) @ V.T) for t in RF.BETA * (x + 1) / 2])
        assert np.max(np.abs(tv - G.tv[m])) < 1e-12
