"""Pinning the equilibrium (Matsubara) GF2 oracle: the reference's stored
G^M of H2 (tests/data/Dsim.h5 with the system of tests/data/hf.hdf5), the
Hartree-Fock closed form, and the closed-form convolution of two free levels."""
import json
import os

import numpy as np

from nedyson_amd import problem
from nedyson_amd.problem import gll_nodes_np

GOLD = os.path.join(os.path.dirname(__file__), "golden")

# Dsim.h5's G^M is sampled on 101 uniform tau points of [0, beta].  The
# inverse temperature and chemical potential are not stored; beta = 50 and
# mu = 0.0423041319 fit the 404 stored values to 4e-10 (mu is the one value
# the fixture fixes only through the tau dependence: the H2 gap makes the
# particle number insensitive to it).
H2_BETA, H2_MU, H2_NTAU = 50.0, 0.0423041319, 100


def h2_reference_gm():
    ds = json.load(open(os.path.join(GOLD, "dsim_invariants.json")))
    return np.array(ds["GM"])[:, 0].reshape(ds["ntau"] + 1, ds["nao"], ds["nao"])


def test_conv_weights_free_levels_closed_form(orc):
    """(g_a o g_b)(tau) = (g_a(tau) - g_b(tau)) / (e_a - e_b) for two free
    fermion levels (the product g_a g_b in Matsubara frequency)."""
    from oracle import matsubara as MS
    beta, ntau = 8.0, 47
    w, tau = MS.conv_weights(ntau, beta)
    ea, eb = 0.7, -0.4
    g = MS.g0_levels([ea, eb], 0.1, tau, beta)
    ga, gb = g[:, 0], g[:, 1]
    conv = np.einsum("mjk,j,k->m", w, ga, gb)
    exact = (ga - gb) / (ea - eb)
    assert np.max(np.abs(conv - exact)) < 1e-12
    # antiperiodicity: a wrong sign on the wrapped piece breaks this by O(1)
    conv_ba = np.einsum("mjk,j,k->m", w, gb, ga)
    assert np.max(np.abs(conv_ba - exact)) < 1e-12


def test_matsubara_oracle_reproduces_reference_h2_gm(orc):
    """The GF2 equilibrium of H2 against the reference's own stored G^M
    (all 2 x 2 x 101 values; tolerance 2e-9)."""
    from oracle import matsubara as MS
    h1, B = MS.h2_system(json.load(open(os.path.join(GOLD, "h2_hf.json"))))
    G, F, mu, it = MS.solve(h1, B, H2_BETA, H2_NTAU, mu=H2_MU)
    L = MS.lagrange(gll_nodes_np(H2_NTAU), np.linspace(-1, 1, H2_NTAU + 1))
    Gu = np.einsum("kj,jab->kab", L, G)
    ref = h2_reference_gm()
    assert np.max(np.abs(Gu - ref)) < 2e-9, np.max(np.abs(Gu - ref))
    # fermionic jump G(0+) + G(beta-) = -1 and 2 tr rho = 2
    assert np.max(np.abs(G[0] + G[-1] + np.eye(2))) < 1e-12
    assert abs(-2 * np.trace(G[-1]) - 2.0) < 1e-9
    # without the self-energy the same system misses the reference by 1e-2
    Ghf, _, _, _ = MS.solve(h1, B, H2_BETA, H2_NTAU, mu=H2_MU, gf2=False)
    assert np.max(np.abs(np.einsum("kj,jab->kab", L, Ghf) - ref)) > 1e-3


def test_matsubara_oracle_hf_limit(orc):
    """Sigma = 0: the self-consistent solution is the Hartree-Fock G^M of the
    closed form (problem.hf_equilibrium + problem.matsubara_gf)."""
    from oracle import matsubara as MS
    p = problem.Problem(nao=4, naux=12, nelec=2, beta=10.0, ntau=31, nt=6, dt=0.02, seed=3, spacing=1.0, b_sd=0.3)
    G, F, mu, it = MS.solve(p.h1, p.B, p.beta, p.ntau, nelec=p.nelec, gf2=False, tol=1e-12)
    assert abs(mu - p.mu) < 1e-8
    assert np.max(np.abs(F - p.fock)) < 1e-8
    assert np.max(np.abs(G - p.Gmat.real)) < 1e-9
