"""Whole runs from the equilibrium GF2 state (ned_matsubara_gf2) on the GPU:
the nao = 4 configuration as specified (sigma_B = 0.3, corrector tolerance
1e-10) against the oracle, total-energy conservation after a smooth switch-on
of the perturbation, the mixed-precision Sigma against FP64 along a
propagated run, and the nao = 32 configuration's window at a history of 500
steps against the oracle."""
import os

import numpy as np
import pytest

import nedyson_amd as ned
from nedyson_amd import problem
from nedyson_amd.green import GreenFunc

pytestmark = pytest.mark.gpu

NTH = min(os.cpu_count() or 1, 16)


def _relerr(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def smoothstep(x):
    """C^5 switch-on from 0 at x <= 0 to 1 at x >= 1 (the first five
    derivatives vanish at both ends)."""
    x = min(max(x, 0.0), 1.0)
    return x ** 6 * (462 - 1980 * x + 3465 * x ** 2 - 3080 * x ** 3 + 1386 * x ** 4 - 252 * x ** 5)


def _equilibrium(p, tol=1e-12):
    with ned.Context(p.nao, p.naux, p.k, p.ntau, p.k) as c:
        c.set_coulomb(p.B)
        mu, F, it, err = c.matsubara_gf2(p.h1, p.beta, nelec=float(p.nelec), mu=p.mu, tol=tol)
        GM = c.read_blocks(ned.NED_G, ned.NED_MAT, 0, p.ntau + 1)
    return mu, F, GM


def _hams(p, ramp):
    """h(T) for T = 0..nt: the abrupt quench h0 from step k+1 on (ramp None),
    or h1 + smoothstep((T - k) dt / ramp) dH."""
    if ramp is None:
        return [p.h0] * (p.nt + 1)
    return [p.h1 + smoothstep((T - p.k) * p.dt / ramp) * p.dH for T in range(p.nt + 1)]


def _gpu_run(p, mu, F, GM, prec, hs, corr_max, corr_tol):
    """Bootstrap and every time step; returns G and per-step (N, E_kin, E_pot)."""
    with ned.Context(p.nao, p.naux, p.nt, p.ntau, p.k) as c:
        c.set_coulomb(p.B)
        c.write_blocks(ned.NED_G, ned.NED_MAT, 0, GM)
        c.set_matsubara(p.beta)
        c.g0_from_h0(p.k, F, mu, p.beta, p.dt)
        try:
            c.do_boot(F, mu, p.beta, p.dt, prec, 60, 1e-12 if prec == "f64" else 1e-8)
        except ned.NedError as e:  # the update stalled above tol: keep the last iterate
            assert e.code == 5, e
            print(f"[boot] {prec} nao={p.nao}: {e}", flush=True)
        obs = []
        for T in range(p.k + 1, p.nt + 1):
            c.do_tstp(T, hs[T], mu, p.beta, p.dt, prec, corr_max, corr_tol)
            e = c.energy(T, hs[T], p.beta, p.dt)
            obs.append((e[2], e[0], e[1]))
        return c.download(ned.NED_G), np.array(obs)


def _oracle_run(orc, p, mu, F, GM, hs, corr_max, corr_tol):
    G = problem.free_gf(F, mu, p.beta, p.dt, p.nt, p.ntau, p.k, GM)
    S = GreenFunc(p.nt, p.ntau, p.nao)
    hmf = np.zeros((p.nt + 1, p.nao, p.nao), dtype=complex)
    D = orc.Dyson(p.nt, p.ntau, p.nao, p.k)
    D.do_boot(p.V, G, S, hmf, F, mu, p.beta, p.dt, 60, 1e-12, nthreads=NTH)
    obs = []
    for T in range(p.k + 1, p.nt + 1):
        D.do_tstp(p.V, T, G, S, hmf, hs[T], mu, p.beta, p.dt, corr_max, corr_tol, nthreads=NTH)
        rho = -1j * G.lesptr(T, T)  # rho = i sig G^<(T,T), sig = -1
        obs.append((2 * np.trace(rho).real, np.sum(rho * (hmf[T] + hs[T]).T).real,
                    2 * D.energy_conv(T, S, G, p.beta, p.dt)))
    return G, np.array(obs)


def test_c0_config_from_gf2_equilibrium(orc):
    """The nao = 4 configuration as specified: sigma_B = 0.3, 64 GLL nodes,
    beta = 10, dt = 0.02, nt = 200, k = 5, the h[0,0] += 0.5 quench, corrector
    tolerance 1e-10, started from the GF2 equilibrium.  The GPU equilibrium
    equals the oracle's; the whole FP64 run equals the oracle's run from the
    same equilibrium on every stored G block and in the energies; the particle
    number is conserved."""
    from oracle import matsubara as MS
    p = problem.config(0)
    mu, F, GM = _equilibrium(p)
    GMo, Fo, muo, _ = MS.solve(p.h1, p.B, p.beta, p.ntau, nelec=p.nelec, tol=1e-12)
    print(f"[c0] equilibrium mu {mu:.10f} (oracle {muo:.10f})", flush=True)
    assert abs(mu - muo) < 1e-9 and np.max(np.abs(GM - GMo)) < 1e-9
    hs = _hams(p, None)
    G, obs = _gpu_run(p, mu, F, GM, "f64", hs, 20, 1e-10)
    Go, obso = _oracle_run(orc, p, mu, F, GM, hs, 20, 1e-10)
    for comp in ("ret", "les", "tv"):
        assert _relerr(getattr(G, comp), getattr(Go, comp)) < 1e-10, comp
    assert np.max(np.abs(obs - obso)) < 1e-10
    print(f"[c0] N drift {np.max(np.abs(obs[:, 0] - obs[0, 0])):.2e}", flush=True)
    assert np.max(np.abs(obs[:, 0] - obs[0, 0])) < 1e-7  # 1.8e-8 across the abrupt quench


def test_energy_conserved_after_smooth_switch_on(orc):
    """GF2 conserves E_kin + E_pot.  With the perturbation switched on smoothly
    (C^5 over 0.5 time units) the k = 5 scheme keeps the total energy constant
    after the switch-on to O(dt^6): nao = 4, sigma_B = 0.3, dt = 0.01, 150
    steps, corrector tolerance 1e-12.  The test is sensitive to the sign and
    size of the potential energy: E_kin - E_pot moves by orders of magnitude
    more than the bound.  The GPU energies equal the oracle's."""
    p = problem.Problem(nao=4, naux=12, nelec=2, beta=10.0, ntau=31, nt=150, dt=0.01, seed=0, spacing=1.0, b_sd=0.3,
                        quench="h00", quench_norm=0.5)
    mu, F, GM = _equilibrium(p)
    hs = _hams(p, 0.5)
    G, obs = _gpu_run(p, mu, F, GM, "f64", hs, 30, 1e-12)
    _, obso = _oracle_run(orc, p, mu, F, GM, hs, 30, 1e-12)
    assert np.max(np.abs(obs - obso)) < 1e-10
    after = np.arange(p.k + 1, p.nt + 1) >= p.k + 50  # the switch-on ends at step k + 50
    E = obs[after, 1] + obs[after, 2]
    dE = np.max(np.abs(E - E[0]))
    swing = np.max(np.abs((obs[after, 1] - obs[after, 2]) - (obs[after, 1] - obs[after, 2])[0]))
    print(f"[energy] drift after switch-on {dE:.2e}, E_kin - E_pot moves {swing:.2e}, "
          f"N drift {np.max(np.abs(obs[:, 0] - obs[0, 0])):.2e}", flush=True)
    assert dE < 1e-7
    assert swing > 1e3 * 1e-7
    assert np.max(np.abs(obs[:, 0] - obs[0, 0])) < 1e-9


@pytest.mark.parametrize("nao,prec", [(4, "f32"), (4, "bf16x3"), (64, "f32"), (64, "bf16x3")])
def test_mixed_precision_run_close_to_f64(nao, prec):
    """Sigma contractions in FP32 / bf16x3, Dyson in FP64, from the GF2
    equilibrium through the quench (3 corrector passes): every stored G block
    within 1e-6 of the FP64 run, and the particle number and total energy
    within 1e-6.  nao = 4 (naux 12, sigma_B 0.1, beta 10, 32 nodes, 60 steps);
    nao = 64 (naux 64, sigma_B 0.05, beta 4, level spacing 0.1, 32 nodes, 20
    steps) takes the batched 64 x 64 chain and the split exchange."""
    if nao == 4:
        p = problem.Problem(nao=4, naux=12, nelec=2, beta=10.0, ntau=31, nt=60, dt=0.01, seed=4, spacing=1.0,
                            b_sd=0.1, quench="h00", quench_norm=0.5)
    else:
        p = problem.Problem(nao=64, naux=64, nelec=32, beta=4.0, ntau=31, nt=20, dt=0.01, seed=64, spacing=0.1,
                            b_sd=0.05, quench="h00", quench_norm=0.5)
    print(f"[mixed] nao={nao} {prec}", flush=True)
    mu, F, GM = _equilibrium(p, tol=1e-10)
    hs = _hams(p, None)
    g64, o64 = _gpu_run(p, mu, F, GM, "f64", hs, 3, 0.0)
    g, o = _gpu_run(p, mu, F, GM, prec, hs, 3, 0.0)
    for comp in ("ret", "les", "tv"):
        err = np.max(np.abs(getattr(g, comp) - getattr(g64, comp)))
        print(f"[mixed] nao={nao} {prec} {comp}: {err:.2e}", flush=True)
        assert err < 1e-6, (comp, err)
    # N = 2 tr rho sums 2 nao diagonal entries, each within the G bound; E
    # within the same bound relative to its size
    dN = np.max(np.abs(o[:, 0] - o64[:, 0]))
    E64 = o64[:, 1] + o64[:, 2]
    dE = np.max(np.abs((o[:, 1] + o[:, 2]) - E64))
    print(f"[mixed] nao={nao} {prec} dN {dN:.2e} dE {dE:.2e} (|E| {np.max(np.abs(E64)):.1f})", flush=True)
    assert dN < 2e-6 * nao
    assert dE < 1e-6 * max(1.0, np.max(np.abs(E64)))


@pytest.mark.timeout(900)
def test_c1_window_T500_parity(orc):
    """The nao = 32 configuration (naux = 128, 96 GLL nodes, dt = 0.02, random
    quench of norm 0.2) at a history of 500 steps: the free G of the Fock
    matrix for t <= 500, Sigma(500) continued time-translation invariantly,
    then one full time step (3 corrector passes) in FP64, GPU against the oracle
    (OpenMP over the Sigma time points), 1e-10.  Covers the long Volterra
    recurrences and multi-chunk history sums at this shape."""
    T0, K = 500, 1
    p = problem.config(1, nt=T0 + K)
    print(f"[c1 window] T0={T0}", flush=True)
    G = problem.free_gf(p.fock, p.mu, p.beta, p.dt, p.nt, p.ntau, T0, p.Gmat)
    S = GreenFunc(p.nt, p.ntau, p.nao)
    hmf = np.zeros((p.nt + 1, p.nao, p.nao), dtype=complex)
    hmf[:T0 + 1] = p.fock
    orc.sigma_solve(p.V, T0, S, G, nthreads=NTH)
    for t in range(T0):  # time-translation-invariant continuation of slice T0
        for tp in range(t + 1):
            S.retptr(t, tp)[:] = S.retptr(T0, T0 - (t - tp))
            S.lesptr(tp, t)[:] = S.lesptr(T0 - (t - tp), T0)
        S.tv[t] = S.tv[T0]
    D = orc.Dyson(p.nt, p.ntau, p.nao, p.k)
    for T in range(T0 + 1, T0 + K + 1):
        D.do_tstp(p.V, T, G, S, hmf, p.h0, p.mu, p.beta, p.dt, 3, nthreads=NTH)
    print("[c1 window] oracle done", flush=True)
    with ned.Context(p.nao, p.naux, p.nt, p.ntau, p.k) as c:
        c.set_coulomb(p.B)
        c.write_blocks(ned.NED_G, ned.NED_MAT, 0, p.Gmat)
        c.set_matsubara(p.beta)
        c.g0_from_h0(T0, p.fock, p.mu, p.beta, p.dt)
        c.write_hmf(np.repeat(p.fock[None].astype(complex), T0 + 1, axis=0))
        c.sigma(T0, "f64")
        c.tti_extend(ned.NED_SIGMA, T0)
        for T in range(T0 + 1, T0 + K + 1):
            c.do_tstp(T, p.h0, p.mu, p.beta, p.dt, "f64", 3)
        s = (T0 + K) * (T0 + K + 1) // 2
        ret = c.read_blocks(ned.NED_G, ned.NED_RET, s, T0 + K + 1)
        les = c.read_blocks(ned.NED_G, ned.NED_LES, s, T0 + K + 1)
        tv = c.read_blocks(ned.NED_G, ned.NED_TV, (T0 + K) * (p.ntau + 1), p.ntau + 1)
    T = T0 + K
    assert _relerr(ret, G.ret_row(T)) < 1e-10
    assert _relerr(les, G.les_col(T)) < 1e-10
    assert _relerr(tv, G.tv[T]) < 1e-10


@pytest.mark.timeout(900)
def test_c2_window_step_parity(orc):
    """The nao = 64 configuration's shape (naux = 256, 128 GLL nodes,
    dt = 0.01; B entries of spread 0.01 as in the benchmark window) at a
    history of 20 steps (free G of the Fock matrix, Sigma(20) continued
    time-translation invariantly), then one full time step of 3 corrector
    passes: FP64 on the GPU against the oracle (1e-10); the same with FP32
    Sigma storage and the G ring (NED_STORE_SIGMA_F32 | NED_STORE_G_RING,
    the storage of the whole nt = 2000 run; 1e-7); and bf16x3 Sigma (1e-6)."""
    T0, K = 20, 1
    p = problem.config(2, nt=T0 + K, b_sd=0.01)
    G = problem.free_gf(p.fock, p.mu, p.beta, p.dt, p.nt, p.ntau, T0, p.Gmat)
    S = GreenFunc(p.nt, p.ntau, p.nao)
    hmf = np.zeros((p.nt + 1, p.nao, p.nao), dtype=complex)
    hmf[:T0 + 1] = p.fock
    print(f"[c2 window] oracle Sigma({T0}) on {NTH} threads", flush=True)
    orc.sigma_solve(p.V, T0, S, G, nthreads=NTH)
    for t in range(T0):
        for tp in range(t + 1):
            S.retptr(t, tp)[:] = S.retptr(T0, T0 - (t - tp))
            S.lesptr(tp, t)[:] = S.lesptr(T0 - (t - tp), T0)
        S.tv[t] = S.tv[T0]
    D = orc.Dyson(p.nt, p.ntau, p.nao, p.k)
    T = T0 + K
    # the step's three corrector passes one call each (the predictor before the
    # first), a progress line after each: a pass takes minutes on the CPU
    for ps in range(3):
        D.do_tstp(p.V, T, G, S, hmf, p.h0, p.mu, p.beta, p.dt, 1, nthreads=NTH, extrapolate=ps == 0)
        print(f"[c2 window] oracle pass {ps + 1} of 3 done", flush=True)
    for storage, prec, tol in ((0, "f64", 1e-10), (ned.NED_STORE_SIGMA_F32 | ned.NED_STORE_G_RING, "f64", 1e-7),
                               (0, "bf16x3", 1e-6)):
        with ned.Context(p.nao, p.naux, p.nt, p.ntau, p.k, storage=storage) as c:
            c.set_coulomb(p.B)
            c.write_blocks(ned.NED_G, ned.NED_MAT, 0, p.Gmat)
            c.set_matsubara(p.beta)
            c.g0_from_h0(T0, p.fock, p.mu, p.beta, p.dt)
            c.write_hmf(np.repeat(p.fock[None].astype(complex), T0 + 1, axis=0))
            c.sigma(T0, "f64")
            c.tti_extend(ned.NED_SIGMA, T0)
            c.do_tstp(T, p.h0, p.mu, p.beta, p.dt, prec, 3)
            s = T * (T + 1) // 2
            ret = c.read_blocks(ned.NED_G, ned.NED_RET, s, T + 1)
            les = c.read_blocks(ned.NED_G, ned.NED_LES, s, T + 1)
            tv = c.read_blocks(ned.NED_G, ned.NED_TV, T * (p.ntau + 1), p.ntau + 1)
        errs = (_relerr(ret, G.ret_row(T)), _relerr(les, G.les_col(T)), _relerr(tv, G.tv[T]))
        print(f"[c2 window] storage {storage} {prec}: relative errors ret/les/tv {errs}", flush=True)
        assert max(errs) < tol, (storage, prec, errs)
