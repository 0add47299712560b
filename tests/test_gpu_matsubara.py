"""The equilibrium GF2 solve on the GPU (ned_matsubara_gf2): against the
reference's stored G^M of H2, against the CPU oracle (with the chemical
potential solved for, and at a size that runs the blocked LU over many
panels), in the Hartree-Fock limit against the closed form, and by
propagating the result without a quench: an equilibrium state must stay
time-translation invariant."""
import json
import os

import numpy as np
import pytest

import nedyson_amd as ned
from nedyson_amd import problem
from nedyson_amd.problem import gll_nodes_np
from oracle import matsubara as MS
from test_matsubara_oracle import GOLD, H2_BETA, H2_MU, H2_NTAU, h2_reference_gm

pytestmark = pytest.mark.gpu


def _solve_gpu(p_nao, B, h1, beta, ntau, nt=6, k=5, **kw):
    c = ned.Context(p_nao, B.shape[0], nt, ntau, k)
    c.set_coulomb(B)
    mu, F, it, err = c.matsubara_gf2(h1, beta, **kw)
    G = c.read_blocks(ned.NED_G, ned.NED_MAT, 0, ntau + 1)
    return c, G, mu, F, it, err


def test_gpu_matsubara_h2_reference_gm():
    """H2 of tests/data/hf.hdf5 at the fixture's beta and mu: all 404 stored
    G^M values within 2e-9, and the oracle within 1e-11."""
    h1, B = MS.h2_system(json.load(open(os.path.join(GOLD, "h2_hf.json"))))
    c, G, mu, F, it, err = _solve_gpu(2, B, h1, H2_BETA, H2_NTAU, mu=H2_MU, tol=1e-12)
    c.close()
    print(f"[matsubara] H2: {it} iterations, last update {err:.2e}", flush=True)
    assert mu == H2_MU
    assert np.max(np.abs(G.imag)) == 0.0
    L = MS.lagrange(gll_nodes_np(H2_NTAU), np.linspace(-1, 1, H2_NTAU + 1))
    dev = np.max(np.abs(np.einsum("kj,jab->kab", L, G.real) - h2_reference_gm()))
    Go, Fo, _, _ = MS.solve(h1, B, H2_BETA, H2_NTAU, mu=H2_MU, tol=1e-13)
    print(f"[matsubara] H2: max |G^M - Dsim| = {dev:.2e}, max |G^M - oracle| = {np.max(np.abs(G.real - Go)):.2e}")
    assert dev < 2e-9
    assert np.max(np.abs(G.real - Go)) < 1e-11
    assert np.max(np.abs(F - Fo)) < 1e-11


def test_gpu_matsubara_vs_oracle_particle_number():
    """nao = 4 GF2 system with mu solved for 2 tr rho = 2: G^M, mu and the
    Fock matrix against the oracle."""
    p = problem.Problem(nao=4, naux=12, nelec=2, beta=10.0, ntau=31, nt=6, dt=0.02, seed=3, spacing=1.0, b_sd=0.3)
    c, G, mu, F, it, err = _solve_gpu(4, p.B, p.h1, p.beta, p.ntau, nelec=2.0, tol=1e-11)
    c.close()
    Go, Fo, muo, ito = MS.solve(p.h1, p.B, p.beta, p.ntau, nelec=2.0, tol=1e-12)
    print(f"[matsubara] nao=4: GPU {it} it, mu {mu:.12f}; oracle {ito} it, mu {muo:.12f}; "
          f"max|dG| {np.max(np.abs(G.real - Go)):.2e}", flush=True)
    assert abs(mu - muo) < 1e-9
    assert np.max(np.abs(G.real - Go)) < 1e-9
    assert np.max(np.abs(F - Fo)) < 1e-9
    assert abs(-2 * np.trace(G[-1].real) - 2.0) < 1e-9


def test_gpu_matsubara_vs_oracle_many_panels():
    """nao = 16, 64 nodes: a 1024 x 1024 system (32 LU panels), mu fixed."""
    p = problem.Problem(nao=16, naux=32, nelec=8, beta=10.0, ntau=63, nt=6, dt=0.02, seed=5, spacing=0.5, b_sd=0.05)
    c, G, mu, F, it, err = _solve_gpu(16, p.B, p.h1, p.beta, p.ntau, mu=p.mu, tol=1e-11)
    c.close()
    Go, Fo, _, ito = MS.solve(p.h1, p.B, p.beta, p.ntau, mu=p.mu, tol=1e-12, nthreads=8)
    print(f"[matsubara] nao=16: GPU {it} it, oracle {ito} it, max|dG| {np.max(np.abs(G.real - Go)):.2e}", flush=True)
    assert np.max(np.abs(G.real - Go)) < 1e-9


def test_gpu_matsubara_hf_limit_closed_form():
    """mode "hf": the self-consistent Hartree-Fock G^M of the closed form."""
    p = problem.Problem(nao=8, naux=24, nelec=4, beta=10.0, ntau=47, nt=6, dt=0.02, seed=7, spacing=0.8, b_sd=0.2)
    c, G, mu, F, it, err = _solve_gpu(8, p.B, p.h1, p.beta, p.ntau, nelec=4.0, mode="hf", tol=1e-12)
    c.close()
    assert abs(mu - p.mu) < 1e-8
    assert np.max(np.abs(F - p.fock)) < 1e-8
    assert np.max(np.abs(G.real - p.Gmat.real)) < 1e-9


def _propagate(c, p, fock, mu, nsteps):
    c.set_matsubara(p.beta)
    c.g0_from_h0(p.k, fock, mu, p.beta, p.dt)
    c.do_boot(fock, mu, p.beta, p.dt, "f64", 60, 1e-12)
    for T in range(p.k + 1, nsteps + 1):
        c.do_tstp(T, p.h1, mu, p.beta, p.dt, "f64", 4, 1e-12)
    ntri = (nsteps + 1) * (nsteps + 2) // 2
    R = c.read_blocks(ned.NED_G, ned.NED_RET, 0, ntri)
    Lz = c.read_blocks(ned.NED_G, ned.NED_LES, 0, ntri)
    s = nsteps // 2
    tti = max(np.max(np.abs(R[(t + s) * (t + s + 1) // 2 + s] - R[t * (t + 1) // 2])) for t in range(nsteps - s + 1))
    dens = [np.trace(Lz[t * (t + 1) // 2 + t]).imag * 2 for t in range(nsteps + 1)]
    return tti, max(abs(d - dens[0]) for d in dens)


def _equilibrium_run(p, nsteps):
    with ned.Context(p.nao, p.naux, p.nt, p.ntau, p.k) as c:
        c.set_coulomb(p.B)
        mu, F, it, err = c.matsubara_gf2(p.h1, p.beta, nelec=float(p.nelec), tol=1e-12)
        return _propagate(c, p, F, mu, nsteps)


def test_gpu_matsubara_equilibrium_is_stationary():
    """Propagating the GF2 equilibrium with h0 = h1 (no quench) keeps
    G^R(t + s, s) = G^R(t, 0) and the density constant (strong coupling,
    sigma_B = 0.3)."""
    nsteps = 40
    p = problem.Problem(nao=4, naux=12, nelec=2, beta=10.0, ntau=31, nt=nsteps, dt=0.02, seed=3, spacing=1.0, b_sd=0.3)
    tti, dn = _equilibrium_run(p, nsteps)
    print(f"[matsubara] stationarity over {nsteps} steps (sigma_B 0.3): tti {tti:.2e} density drift {dn:.2e}",
          flush=True)
    assert tti < 1e-6 and dn < 1e-6


def test_gpu_matsubara_stationarity_detects_non_equilibrium():
    """The same check is sensitive: at sigma_B = 0.1 the Hartree-Fock state
    (not an equilibrium of GF2) breaks time-translation invariance by orders
    of magnitude more than the GF2 equilibrium."""
    nsteps = 40
    p = problem.Problem(nao=4, naux=12, nelec=2, beta=10.0, ntau=31, nt=nsteps, dt=0.02, seed=3, spacing=1.0, b_sd=0.1)
    tti, dn = _equilibrium_run(p, nsteps)
    with ned.Context(p.nao, p.naux, p.nt, p.ntau, p.k) as c:
        c.set_coulomb(p.B)
        c.upload(ned.NED_G, p.initial_green())
        tti_hf, dn_hf = _propagate(c, p, p.fock, p.mu, nsteps)
    print(f"[matsubara] sigma_B 0.1: GF2 equilibrium tti {tti:.2e} drift {dn:.2e}; "
          f"HF start tti {tti_hf:.2e} drift {dn_hf:.2e}", flush=True)
    assert tti < 1e-9 and dn < 1e-9
    assert tti_hf > 1e4 * tti


def test_gpu_replays_reference_h2_propagation(orc):
    """End to end on the GPU: ned_matsubara_gf2, bootstrap and 95 FP64 time
    steps of H2 against tests/data/Dsim.h5 (the same bounds as the oracle's
    replay), and against the oracle's replay to 1e-11."""
    from test_matsubara_oracle import H2_DT, H2_K, H2_NT, check_h2_replay, h2_replay_oracle
    h1, B = MS.h2_system(json.load(open(os.path.join(GOLD, "h2_hf.json"))))
    with ned.Context(2, B.shape[0], H2_NT, H2_NTAU, H2_K) as c:
        c.set_coulomb(B)
        mu, F, _, _ = c.matsubara_gf2(h1, H2_BETA, mu=H2_MU, tol=1e-13)
        c.set_matsubara(H2_BETA)
        c.g0_from_h0(H2_K, F, mu, H2_BETA, H2_DT)
        c.do_boot(F, mu, H2_BETA, H2_DT, "f64", 60, 1e-13)
        for T in range(H2_K + 1, H2_NT + 1):
            c.do_tstp(T, h1, mu, H2_BETA, H2_DT, "f64", 4, 1e-13)
        ntri = (H2_NT + 1) * (H2_NT + 2) // 2
        R = c.read_blocks(ned.NED_G, ned.NED_RET, 0, ntri)
        Lz = c.read_blocks(ned.NED_G, ned.NED_LES, 0, ntri)
    GR0 = np.array([R[t * (t + 1) // 2] for t in range(H2_NT + 1)])
    GLtt = np.array([Lz[t * (t + 1) // 2 + t] for t in range(H2_NT + 1)])
    dR, dL = check_h2_replay(GR0, GLtt)
    oR, oL, _ = h2_replay_oracle(orc)
    assert np.max(np.abs(GR0 - oR)) < 1e-11 and np.max(np.abs(GLtt - oL)) < 1e-11
    print(f"[matsubara] H2 replay: max|G^R(t,0) - Dsim| t<=10 {dR[:11].max():.2e}, all t {dR.max():.2e}", flush=True)
