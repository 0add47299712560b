"""CPU oracle of the equilibrium (Matsubara) GF2 solve -- test infrastructure
only; nothing in the product imports it.

The reference takes its initial state from gfmol (gfmol::DecompSimulation,
built at src/molNEsim_decomp_impl.h:21, run at :50-52, sampled on the
collocation nodes at :145-165).  gfmol is an external project with no pinned
version, so this restates the equations it solves:

    G = G0 + G0 o Sigma o G,   (A o B)(tau) = int_0^beta A(tau - t) B(t) dt,

A(s < 0) = -A(s + beta) (fermions), G0 = G0[F - mu] with the Fock matrix
F = h1 + 2J[rho] - K[rho], rho = -G(beta), and Sigma the second-Born
self-energy, obtained from the real-time GF2 oracle on the T = 0 slice:
G^tv(0, tau_m) = -i G(beta - tau_m) and Sigma(beta - tau_m) = i Sigma^tv(0, tau_m).

Discretisation: nodal interpolants on the ntau+1 Gauss-Lobatto-Legendre
nodes, convolutions integrated exactly (Gauss-Legendre on the two
polynomial pieces), and the linear Dyson system solved densely with numpy.
mu is found by bisection on 2 tr rho = nelec; the outer loop is plain damped
iteration with Anderson mixing.  The result is pinned by the reference's
fixture tests/data/Dsim.h5 (the G^M of H2, tests/data/hf.hdf5).
"""
import numpy as np

from nedyson_amd.green import GreenFunc
from nedyson_amd.problem import fermi, gll_nodes_np, hf_fock

from . import oracle as O


def lagrange(xn, x):
    """L[k, j] = l_j(x[k]) for the Lagrange basis on the nodes xn."""
    xn = np.asarray(xn, dtype=float)
    bw = np.array([1.0 / np.prod(np.delete(xn[j] - xn, j)) for j in range(len(xn))])
    out = np.zeros((len(x), len(xn)))
    for k, xv in enumerate(np.atleast_1d(x)):
        d = xv - xn
        hit = np.flatnonzero(d == 0)
        if hit.size:
            out[k, hit[0]] = 1.0
            continue
        t = bw / d
        out[k] = t / t.sum()
    return out


def conv_weights(ntau, beta, sig=-1):
    """w[m, j, k]: C(tau_m) = sum_{j,k} w[m,j,k] A_j B_k for the antiperiodic
    convolution of the nodal interpolants (exact)."""
    x = gll_nodes_np(ntau)
    tau = beta * (x + 1) / 2
    npt = ntau + 1
    y, W = np.polynomial.legendre.leggauss(npt)
    w = np.zeros((npt, npt, npt))
    for m, tm in enumerate(tau):
        for lo, hi, fac, shift in ((0.0, tm, 1.0, 0.0), (tm, beta, float(sig), beta)):
            if hi <= lo:
                continue
            t = lo + (hi - lo) * (y + 1) / 2
            wq = fac * W * (hi - lo) / 2
            La = lagrange(x, 2 * (tm - t + shift) / beta - 1)
            Lb = lagrange(x, 2 * t / beta - 1)
            w[m] += np.einsum("q,qj,qk->jk", wq, La, Lb)
    return w, tau


def g0_levels(eps, mu, tau, beta):
    """g[m, p]: the Matsubara function of a level eps_p at tau_m."""
    lam = mu - np.asarray(eps)[None, :]
    t = np.asarray(tau)[:, None]
    return -np.where(lam < 0, np.exp(lam * t) * fermi(beta, lam), np.exp((t - beta) * lam) * fermi(beta, -lam))


def sigma_matsubara(V, GM, ntau, nthreads=1):
    """Sigma^M on the nodes from the real-time GF2 oracle at T = 0."""
    n = GM.shape[1]
    G, S = GreenFunc(0, ntau, n), GreenFunc(0, ntau, n)
    for m in range(ntau + 1):
        G.tv[0, m] = -1j * GM[ntau - m]
    G.les[0] = G.tv[0, 0]
    G.ret[0] = -1j * np.eye(n)
    O.sigma_solve(V, 0, S, G, nthreads=nthreads)
    SM = np.zeros((ntau + 1, n, n))
    for m in range(ntau + 1):
        SM[ntau - m] = (1j * S.tv[0, m]).real
    return SM


def dyson(w, tau, beta, F, mu, SM):
    """Solve G = G0[F - mu] + G0 o SM o G for G (real)."""
    npt, n = len(tau), F.shape[0]
    eps, U = np.linalg.eigh(F)
    g = g0_levels(eps, mu, tau, beta)                     # (npt, n)
    Sp = np.einsum("ap,jab,bq->jpq", U, SM, U)
    K = np.einsum("mjk,jpq->mpkq", w, Sp).reshape(npt * n, npt * n)
    Cb = np.einsum("mjk,jp->pmk", w, g)                   # block of orbital p
    K4 = K.reshape(npt, n, npt * n)
    M = np.eye(npt * n) - np.einsum("pmk,kpx->mpx", Cb, K4).reshape(npt * n, npt * n)
    rhs = np.zeros((npt, n, n))
    rhs[:, np.arange(n), np.arange(n)] = g
    Gp = np.linalg.solve(M, rhs.reshape(npt * n, n)).reshape(npt, n, n)
    return np.einsum("ap,mpq,bq->mab", U, Gp, U)


def solve(h1, B, beta, ntau, nelec=None, mu=None, gf2=True, tol=1e-11, maxiter=200, hist=6, G=None, nthreads=1):
    """Self-consistent equilibrium.  nelec given: mu by bisection (mu is
    then only a starting point); otherwise mu fixed.  Returns (G, F, mu, iters)."""
    V = np.ascontiguousarray(np.transpose(B, (1, 2, 0)))
    w, tau = conv_weights(ntau, beta)
    n = h1.shape[0]
    if G is None:
        eps, U = np.linalg.eigh(h1)
        if mu is None:
            occ = int(round(nelec / 2))
            mu = 0.5 * (eps[max(occ - 1, 0)] + eps[min(occ, n - 1)])
        G = np.einsum("ap,mp,bp->mab", U, g0_levels(eps, mu, tau, beta), U)
    Xs, Rs = [], []
    for it in range(maxiter):
        SM = sigma_matsubara(V, G, ntau, nthreads) if gf2 else np.zeros_like(G)
        F = hf_fock(h1, B, -G[-1])
        if nelec is not None:
            def count(m):
                return -2 * np.trace(dyson(w, tau, beta, F, m, SM)[-1])
            lo, hi = mu - 0.5, mu + 0.5
            while count(lo) > nelec:
                lo -= 2 * (hi - lo)
            while count(hi) < nelec:
                hi += 2 * (hi - lo)
            for _ in range(100):
                mid = 0.5 * (lo + hi)
                if count(mid) > nelec:
                    hi = mid
                else:
                    lo = mid
                if hi - lo < 1e-14 * max(1.0, abs(mid)):
                    break
            mu = 0.5 * (lo + hi)
        Gn = dyson(w, tau, beta, F, mu, SM)
        r = Gn - G
        if np.max(np.abs(r)) < tol:
            return Gn, hf_fock(h1, B, -Gn[-1]), mu, it + 1
        Xs.append(G.copy())
        Rs.append(r.copy())
        Xs, Rs = Xs[-hist:], Rs[-hist:]
        if len(Xs) > 1:
            dR = np.array([(Rs[i + 1] - Rs[i]).ravel() for i in range(len(Xs) - 1)])
            dX = np.array([(Xs[i + 1] - Xs[i]).ravel() for i in range(len(Xs) - 1)])
            gam = np.linalg.lstsq(dR.T, r.ravel(), rcond=None)[0]
            G = (G.ravel() + r.ravel() - (dX + dR).T @ gam).reshape(G.shape)
        else:
            G = G + r
    raise RuntimeError("matsubara oracle did not converge")


def h2_system(d):
    """(h1, B) of the tests/data/hf.hdf5 H2 system: B from the eigen-
    decomposition of the (ij|kl) matrix, symmetric in (i, j)."""
    n = d["nao"]
    h1 = np.array(d["hcore"]).reshape(n, n)
    U = np.array(d["uchem"]).reshape(n * n, n * n)
    e, v = np.linalg.eigh(U)
    keep = e > 1e-12 * e.max()
    B = (v[:, keep] * np.sqrt(e[keep])).T.reshape(-1, n, n)
    return h1, 0.5 * (B + B.transpose(0, 2, 1))
