"""Seeded synthetic systems and their initial (equilibrium) state.

The reference reads h1 and the Cholesky-factored Coulomb tensor from a pyscf
HDF5 file and gets the equilibrium state from the external gfmol solver
(src/molNEsim_decomp_impl.h:14-52).  Neither is available here, so a problem
is generated from a seed:

* ``h1``: real symmetric; diagonal ``spacing * (i - (nao-1)/2)`` plus
  ``0.1*spacing`` times sorted N(0,1) draws, off-diagonals N(0, 0.1)
  symmetrised.
* ``B[Q,i,j]``: upper triangle N(0, sigma_B), mirrored (symmetric in ij).
* equilibrium: restricted Hartree-Fock at inverse temperature beta with
  ``2 tr rho = nelec``; mu by bisection.  G^M is the Matsubara function of the
  HF Fock matrix on the Gauss-Lobatto-Legendre nodes (src/dyson_free.h:
  161-176 form).  The GF2 correlation of the initial state is therefore left
  out (the gfmol Matsubara GF2 solver is not part of this build).
* quench: ``h0 = h1 + dH`` (post-quench one-body Hamiltonian).

Everything here runs once on the host to produce inputs shared by the GPU
path and the CPU oracle.
"""
import numpy as np


def gll_nodes_np(ntau):
    """GLL nodes on [-1,1] (ascending), Newton on (1-x^2) P_N'(x)."""
    N = ntau
    x = -np.cos(np.pi * np.arange(N + 1) / N)
    for _ in range(100):
        p0 = np.ones_like(x)
        p1 = x.copy()
        for l in range(2, N + 1):
            p0, p1 = p1, ((2 * l - 1) * x * p1 - (l - 1) * p0) / l
        dx = (x * p1 - p0) / ((N + 1) * p1)
        x = x - dx
        if np.max(np.abs(dx)) < 1e-16:
            break
    x[0], x[-1] = -1.0, 1.0
    return np.sort(x)


def fermi(beta, w):
    a = np.clip(beta * np.asarray(w), -700, 700)
    return 1.0 / (1.0 + np.exp(a))


def make_h1(rng, nao, spacing, off_sd=0.1):
    d = spacing * (np.arange(nao) - (nao - 1) / 2.0) + 0.1 * spacing * np.sort(rng.normal(0, 1, nao))
    h = np.triu(rng.normal(0, off_sd, (nao, nao)), 1)
    return h + h.T + np.diag(d)


def make_B(rng, nao, naux, sd):
    B = np.zeros((naux, nao, nao))
    iu = np.triu_indices(nao)
    for q in range(naux):
        v = rng.normal(0, sd, len(iu[0]))
        B[q][iu] = v
        B[q] = B[q] + np.triu(B[q], 1).T
    return B


def hf_fock(h, B, rho):
    """h + 2J[rho] - K[rho] with V(i,j,Q) = B[Q,i,j] (src/molHFSolver.cpp:81-110)."""
    X = np.einsum("kl,qkl->q", rho, B)
    J = np.einsum("q,qij->ij", X, B)
    K = np.sum(np.matmul(np.matmul(B, rho), B), axis=0)
    return h + 2 * J - K


def density(F, beta, nelec):
    """rho = f(F - mu), mu bisected so that 2 tr rho = nelec."""
    e, v = np.linalg.eigh(F)
    lo, hi = e[0] - 50.0 / beta - 10, e[-1] + 50.0 / beta + 10
    for _ in range(200):
        mu = 0.5 * (lo + hi)
        n = 2 * fermi(beta, e - mu).sum()
        if n > nelec:
            hi = mu
        else:
            lo = mu
    mu = 0.5 * (lo + hi)
    occ = fermi(beta, e - mu)
    return (v * occ) @ v.T, mu


def _anderson_fock(h, B, beta, nelec, Fin, tol, maxiter, hist=8):
    """Anderson (Pulay) mixing of the Fock matrix: F_in -> rho = f(F_in - mu)
    -> F_out = h + 2J - K, minimising the residual over the last iterates."""
    ins, outs = [], []
    r = np.inf
    for _ in range(maxiter):
        rho, mu = density(Fin, beta, nelec)
        Fout = hf_fock(h, B, rho)
        r = np.max(np.abs(Fout - Fin))
        if r < tol:
            break
        ins.append(Fin)
        outs.append(Fout)
        ins, outs = ins[-hist:], outs[-hist:]
        m = len(ins)
        R = np.array([(o - i).ravel() for i, o in zip(ins, outs)])
        A = np.zeros((m + 1, m + 1))
        A[:m, :m] = R @ R.T
        A[:m, m] = A[m, :m] = 1.0
        rhs = np.zeros(m + 1)
        rhs[m] = 1.0
        c = np.linalg.lstsq(A, rhs, rcond=None)[0][:m]
        Fin = 0.5 * (sum(ci * x for ci, x in zip(c, ins)) + sum(ci * x for ci, x in zip(c, outs)))
    return Fin, r


def hf_equilibrium(h, B, beta, nelec, tol=1e-10, accept=1e-8):
    """Finite-temperature restricted HF.  Direct Anderson iteration at beta; if
    that stalls, anneal from high temperature, reusing each solution as the
    start of the next.  Returns (Fock, rho, mu); the final residual
    max|F_out - F_in| is stored in hf_equilibrium.residual."""
    F0 = hf_fock(h, B, density(h, beta, nelec)[0])
    F, r = _anderson_fock(h, B, beta, nelec, F0, tol, 300)
    if r >= tol:
        F = hf_fock(h, B, density(h, min(beta, 0.5), nelec)[0])
        for b in [x for x in (0.5, 1, 2, 4, 8, 12, 16) if x < beta] + [beta]:
            F, r = _anderson_fock(h, B, b, nelec, F, tol, 400)
    if r >= accept:
        raise RuntimeError(f"HF equilibrium did not converge (residual {r:.2e})")
    hf_equilibrium.residual = r
    rho, mu = density(F, beta, nelec)
    F = hf_fock(h, B, rho)
    rho, mu = density(F, beta, nelec)
    return F, rho, mu


def matsubara_gf(F, mu, beta, ntau):
    """G^M(tau_m) = -V diag(e^{lam tau}/(1+e^{lam beta})) V^T, lam = eig(mu - F)."""
    x = gll_nodes_np(ntau)
    taus = beta * (x + 1) / 2
    lam, V = np.linalg.eigh(mu * np.eye(F.shape[0]) - F)
    G = np.zeros((ntau + 1,) + F.shape, dtype=np.complex128)
    for m, t in enumerate(taus):
        fe = np.where(lam < 0, np.exp(lam * t) * fermi(beta, lam), np.exp((t - beta) * lam) * fermi(beta, -lam))
        G[m] = -(V * fe) @ V.T
    return G


def free_gf(fock, mu, beta, dt, nt, ntau, k, Gmat=None):
    """Initial guess of the bootstrap: G0_from_h0 with the equilibrium Fock
    matrix (src/dyson_free.h:226-297, called at src/molNEsim_decomp_impl.h:44-46)
    for steps 0..k, and G^M on the GLL nodes."""
    from .green import GreenFunc
    nao = fock.shape[0]
    lam, V = np.linalg.eigh(mu * np.eye(nao) - fock)
    taus = beta * (gll_nodes_np(ntau) + 1) / 2
    g = GreenFunc(nt, ntau, nao, -1)
    g.mat[:] = matsubara_gf(fock, mu, beta, ntau) if Gmat is None else Gmat

    def fexp(tau, w):
        return np.where(w < 0, np.exp(w * tau) * fermi(beta, w), np.exp((tau - beta) * w) * fermi(beta, -w))

    U = [V @ np.diag(np.exp(1j * dt * n * lam)) @ V.T for n in range(k + 1)]
    occ = V @ np.diag(fermi(beta, -lam)) @ V.T
    for n in range(k + 1):
        for m, t in enumerate(taus):
            g.tv[n, m] = 1j * U[n] @ ((V * fexp(t, -lam)) @ V.T)
    for m in range(k + 1):
        for n in range(m + 1):
            g.retptr(m, n)[:] = -1j * U[m] @ U[n].conj().T
            g.lesptr(n, m)[:] = 1j * U[n] @ occ @ U[m].conj().T
    return g


class Problem:
    """A seeded KBE-GF2 system (see module docstring)."""

    def __init__(self, nao, naux, nelec, beta, ntau, nt, dt, k=5, seed=0, spacing=1.0, b_sd=0.3,
                 quench="h00", quench_norm=0.5):
        rng = np.random.default_rng(seed)
        self.nao, self.naux, self.nelec, self.beta = nao, naux, nelec, beta
        self.ntau, self.nt, self.dt, self.k, self.seed = ntau, nt, dt, k, seed
        self.b_sd, self.quench, self.quench_norm = b_sd, quench, quench_norm
        self.h1 = make_h1(rng, nao, spacing)
        self.B = make_B(rng, nao, naux, b_sd)
        self.V = np.ascontiguousarray(np.transpose(self.B, (1, 2, 0)))  # Vija(i,j,Q)
        self.fock, self.rho_eq, self.mu = hf_equilibrium(self.h1, self.B, beta, nelec)
        self.Gmat = matsubara_gf(self.fock, self.mu, beta, ntau)
        if quench == "h00":
            dH = np.zeros((nao, nao))
            dH[0, 0] = quench_norm
        elif quench == "random":
            dH = rng.normal(0, 1, (nao, nao))
            dH = dH + dH.T
            dH *= quench_norm / np.linalg.norm(dH)
        else:
            dH = np.zeros((nao, nao))
        self.dH = dH
        self.h0 = self.h1 + dH

    def initial_green(self):
        return free_gf(self.fock, self.mu, self.beta, self.dt, self.nt, self.ntau, self.k, self.Gmat)


BATCH_NAOS = (8, 12, 16, 24, 32, 48)


def batch_systems(nsys=16, seed=3000, **over):
    """The batched configuration: nsys independent systems, nao drawn
    uniformly from BATCH_NAOS (seeded), naux = 4 nao, nao/2 electrons,
    beta = 10, 64 GLL nodes, dt = 0.02, nt = 1500, k = 5, the same generators
    (level spacing 0.3), each with its own random quench of norm drawn
    uniformly from [0.1, 0.3].  over overrides any Problem argument (b_sd, nt,
    ...) for every system."""
    rng = np.random.default_rng(seed)
    naos = rng.choice(BATCH_NAOS, size=nsys)
    norms = rng.uniform(0.1, 0.3, size=nsys)
    out = []
    for i, (n, qn) in enumerate(zip(naos, norms)):
        p = dict(nao=int(n), naux=4 * int(n), nelec=int(n) // 2, beta=10.0, ntau=63, nt=1500, dt=0.02, k=5,
                 seed=seed + 1 + i, spacing=0.3, b_sd=0.1, quench="random", quench_norm=float(qn))
        p.update(over)
        out.append(Problem(**p))
    return out


PREDICTOR_NSYS = 64  # systems 0..47 train the network, 48..63 test it


def predictor_system(i, seed=4000, **over):
    """System i of the learned-predictor configuration: 64 independent nao = 48
    systems (naux = 192, 24 electrons, beta = 20, 96 GLL nodes, dt = 0.02,
    nt = 4000, k = 5, the same generators), each with its own random quench of
    norm drawn uniformly from [0.05, 0.4]; systems 0..47 train the network,
    48..63 test it.  over overrides any Problem argument."""
    if not 0 <= i < PREDICTOR_NSYS:
        raise ValueError(i)
    norms = np.random.default_rng(seed).uniform(0.05, 0.4, size=PREDICTOR_NSYS)
    p = dict(nao=48, naux=192, nelec=24, beta=20.0, ntau=95, nt=4000, dt=0.02, k=5, seed=seed + 1 + i,
             spacing=0.3, b_sd=0.1, quench="random", quench_norm=float(norms[i]))
    p.update(over)
    return Problem(**p)


def config(idx, **over):
    """The benchmark configurations c0..c2 (spin-restricted GF2, k=5)."""
    if idx == 0:
        p = dict(nao=4, naux=12, nelec=2, beta=10.0, ntau=63, nt=200, dt=0.02, k=5, seed=0, spacing=1.0, b_sd=0.3,
                 quench="h00", quench_norm=0.5)
    elif idx == 1:
        p = dict(nao=32, naux=128, nelec=16, beta=20.0, ntau=95, nt=1000, dt=0.02, k=5, seed=1000, spacing=0.3,
                 b_sd=0.1, quench="random", quench_norm=0.2)
    elif idx == 2:
        p = dict(nao=64, naux=256, nelec=32, beta=20.0, ntau=127, nt=2000, dt=0.01, k=5, seed=2000, spacing=0.3,
                 b_sd=0.1, quench="random", quench_norm=0.2)
    else:
        raise ValueError(idx)
    p.update(over)
    return Problem(**p)
