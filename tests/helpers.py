"""Shared test inputs: random contour functions, the GF2 definition in numpy,
exact free propagation, polynomial contour data."""
import numpy as np

from nedyson_amd.green import GreenFunc


def random_green(rng, nt, ntau, nao, scale=1.0):
    """Uniform random entries in every stored block (as Eigen::Random in
    tests/mol2bSolver-test.cpp:106-108)."""
    g = GreenFunc(nt, ntau, nao, -1)
    for a in (g.ret, g.les, g.tv, g.mat):
        a[:] = scale * (rng.uniform(-1, 1, a.shape) + 1j * rng.uniform(-1, 1, a.shape))
    return g


def sym_V(rng, nao, naux, sd=0.3):
    B = rng.normal(0, sd, (naux, nao, nao))
    B = 0.5 * (B + np.transpose(B, (0, 2, 1)))
    return np.ascontiguousarray(np.transpose(B, (1, 2, 0))), B


def gf2_functional(V, A, Bm):
    """F[A,B]_ij = sum (2 U_{il,np} U_{kj,qm} - U_{il,np} U_{qj,km}) A_lk B_mn A_pq,
    the index pattern of solve_loop (src/mol2bSolver_decomp_impl.h:459-467, 585-593),
    written directly with the full U = V V^T."""
    U = np.einsum("ijq,klq->ijkl", V, V)
    bub = 2 * np.einsum("ilnp,kjqm,lk,mn,pq->ij", U, U, A, Bm, A, optimize=True)
    exc = np.einsum("ilnp,qjkm,lk,mn,pq->ij", U, U, A, Bm, A, optimize=True)
    return bub - exc


def free_green(h, mu, beta, dt, nt, x, sig=-1):
    """Exact non-interacting G for constant h (src/dyson_free.h:222-297 closed forms)
    for all stored times."""
    nao = h.shape[0]
    lam, V = np.linalg.eigh(mu * np.eye(nao) - h)
    ntau = len(x) - 1
    taus = beta * (np.asarray(x) + 1) / 2
    g = GreenFunc(nt, ntau, nao, sig)

    def fermi(w):
        return 1 / (1 + np.exp(np.clip(beta * w, -700, 700)))

    def fexp(tau, w):
        return np.where(w < 0, np.exp(w * tau) * fermi(w), np.exp((tau - beta) * w) * fermi(-w))

    U = [V @ np.diag(np.exp(1j * dt * n * lam)) @ V.T for n in range(nt + 1)]
    occ = V @ np.diag(fermi(-lam)) @ V.T
    for m in range(ntau + 1):
        g.mat[m] = -(V * fexp(taus[m], lam)) @ V.T
    F = [V @ np.diag(fexp(t, -lam)) @ V.T for t in taus]
    for n in range(nt + 1):
        for m in range(ntau + 1):
            g.tv[n, m] = 1j * U[n] @ F[m]
    for m in range(nt + 1):
        for n in range(m + 1):
            g.retptr(m, n)[:] = -1j * U[m] @ U[n].conj().T
            g.lesptr(n, m)[:] = 1j * U[n] @ occ @ U[m].conj().T
    return g


def poly_green(nt, ntau, nao, dt, deg, rng):
    """Contour data polynomial of degree <= deg in the times, with the
    symmetries the predictor relies on: G^R(t,t') = i p(t,t') M,
    G^<(t,t') = i q(t,t') M' (p, q symmetric), G^rceil(t,0) = -G^<(0,t)^H."""
    M = rng.normal(size=(nao, nao)) + 1j * rng.normal(size=(nao, nao))
    M = M + M.conj().T
    Mp = rng.normal(size=(nao, nao)) + 1j * rng.normal(size=(nao, nao))
    Mp = Mp + Mp.conj().T
    cp = rng.normal(size=deg + 1)
    cq = rng.normal(size=deg + 1)

    def p(t, s):
        return sum(cp[a] * ((t + s) * dt) ** a for a in range(deg + 1)) + 0.3 * (t * s * dt * dt) ** (deg // 2)

    def q(t, s):
        return sum(cq[a] * ((t + s) * dt) ** a for a in range(deg + 1))

    g = GreenFunc(nt, ntau, nao, -1)
    for t in range(nt + 1):
        for s in range(t + 1):
            g.retptr(t, s)[:] = 1j * p(t, s) * M
            g.lesptr(s, t)[:] = 1j * q(s, t) * Mp
        for m in range(ntau + 1):
            g.tv[t, m] = 1j * q(0, t) * Mp + m * (t * dt) ** min(deg, 2) * M
    return g
