"""Independent systems of uneven size stepped together (ned_batch_do_tstp):
the same G as stepping each system alone, FP32 Sigma and FP64 Dyson.
(B entries drawn with sigma 0.05: at 0.1 the nao = 12 system started from the
Hartree-Fock state diverges within 25 steps, on the CPU oracle as well.)"""
import time

import numpy as np
import pytest

import nedyson_amd as ned
from nedyson_amd import problem

pytestmark = pytest.mark.gpu

NAOS = (4, 6, 8, 12)


def _systems(nt):
    rng = np.random.default_rng(3000)
    out = []
    for i, n in enumerate(NAOS):
        out.append(problem.Problem(nao=n, naux=4 * n, nelec=n // 2, beta=10.0, ntau=15, nt=nt, dt=0.02,
                                   seed=3000 + i, spacing=0.3, b_sd=0.05, quench="random",
                                   quench_norm=float(rng.uniform(0.1, 0.3))))
    return out


def _booted(p):
    c = ned.Context(p.nao, p.naux, p.nt, p.ntau, p.k)
    c.set_coulomb(p.B)
    c.upload(ned.NED_G, p.initial_green())
    c.set_matsubara(p.beta)
    c.do_boot(p.fock, p.mu, p.beta, p.dt, "f64", 50, 1e-10)
    return c


def test_batch_matches_independent_runs():
    nt = 30
    ps = _systems(nt)
    alone, t_alone = [], 0.0
    for p in ps:
        with _booted(p) as c:
            t0 = time.perf_counter()
            for T in range(p.k + 1, nt + 1):
                c.do_tstp(T, p.h0, p.mu, p.beta, p.dt, "f32", 3)
            c.sync()
            t_alone += time.perf_counter() - t0
            alone.append(c.download(ned.NED_G))
    ctxs = [_booted(p) for p in ps]
    b = ned.Batch(ctxs)
    t0 = time.perf_counter()
    for T in range(ps[0].k + 1, nt + 1):
        it, err = b.do_tstp(T, [p.h0 for p in ps], [p.mu for p in ps], [p.beta for p in ps], [p.dt for p in ps],
                            "f32", 3)
        assert it == [3] * len(ps) and np.all(np.isfinite(err))
    for c in ctxs:
        c.sync()
    t_batch = time.perf_counter() - t0
    print(f"[batch] nao {NAOS}: {nt - ps[0].k} steps alone one after another {t_alone:.2f} s, "
          f"together {t_batch:.2f} s", flush=True)
    for p, c, ref in zip(ps, ctxs, alone):
        d = c.download(ned.NED_G)
        for comp in ("ret", "les", "tv"):
            assert np.max(np.abs(getattr(d, comp) - getattr(ref, comp))) <= 1e-12, (p.nao, comp)
    b.close()
    for c in ctxs:
        c.close()


def test_batch_reports_the_failing_system():
    """A system whose step cannot run (tstp beyond its nt) fails alone; the
    others are stepped and the status names it."""
    ps = _systems(10)
    ctxs = [_booted(p) for p in ps]
    short = ned.Context(4, 16, 5, 15, 5)  # nt = 5 < tstp
    short.set_coulomb(ps[0].B)
    short.upload(ned.NED_G, problem.free_gf(ps[0].fock, ps[0].mu, ps[0].beta, ps[0].dt, 5, 15, 5, ps[0].Gmat))
    short.set_matsubara(ps[0].beta)
    b = ned.Batch(ctxs + [short])
    with pytest.raises(ned.NedError) as e:
        b.do_tstp(6, [p.h0 for p in ps] + [ps[0].h0], [p.mu for p in ps] + [ps[0].mu],
                  [p.beta for p in ps] + [ps[0].beta], [p.dt for p in ps] + [ps[0].dt], "f32", 3)
    assert "system 4" in str(e.value)
    b.close()
    for c in ctxs + [short]:
        c.close()


def test_batch_threads_mode_matches_grouped(monkeypatch):
    """The two batch modes (grouped Dyson launches, and one host thread and
    stream per system) give the same G.  The grouped recurrences launch one
    dependency-chain link of two products at a time, so the contexts are made
    to do the same (NED_CHAIN=0) for a bitwise comparison."""
    monkeypatch.setenv("NED_CHAIN", "0")
    nt = 16
    ps = _systems(nt)
    outs = []
    for grouped in (True, False):
        ctxs = [_booted(p) for p in ps]
        b = ned.Batch(ctxs, grouped=grouped)
        for T in range(ps[0].k + 1, nt + 1):
            b.do_tstp(T, [p.h0 for p in ps], [p.mu for p in ps], [p.beta for p in ps], [p.dt for p in ps], "f32", 3)
        outs.append([c.download(ned.NED_G) for c in ctxs])
        b.close()
        for c in ctxs:
            c.close()
    for a, b_ in zip(*outs):
        for comp in ("ret", "les", "tv"):
            assert np.array_equal(getattr(a, comp), getattr(b_, comp)), comp


def _equilibrium(p):
    with ned.Context(p.nao, p.naux, p.k, p.ntau, p.k) as c:
        c.set_coulomb(p.B)
        mu, F, it, err = c.matsubara_gf2(p.h1, p.beta, nelec=float(p.nelec), mu=p.mu, tol=1e-9)
        return mu, F, c.read_blocks(ned.NED_G, ned.NED_MAT, 0, p.ntau + 1)


def _equilibrium_ctx(p, nt, mu, F, GM):
    c = ned.Context(p.nao, p.naux, nt, p.ntau, p.k)
    c.set_coulomb(p.B)
    c.write_blocks(ned.NED_G, ned.NED_MAT, 0, GM)
    c.set_matsubara(p.beta)
    c.g0_from_h0(p.k, F, mu, p.beta, p.dt)
    try:
        c.do_boot(F, mu, p.beta, p.dt, "f32", 60, 1e-9)
    except ned.NedError as e:  # the update stalled above tol: keep the last iterate
        assert e.code == 5, e
    return c


@pytest.mark.timeout(900)
def test_batch_configuration_16_uneven_systems(orc, monkeypatch):
    """The batched configuration's systems (problem.batch_systems: 16 systems,
    nao drawn from {8, 12, 16, 24, 32, 48}, naux = 4 nao, 64 GLL nodes,
    dt = 0.02, each with its own quench), B entries with sigma 0.05, each from
    its GF2 equilibrium through 12 steps of FP32 Sigma and FP64 Dyson: the
    grouped batch equals the systems stepped alone, bit for bit, and the
    nao = 8 system equals the oracle's FP64 run from the same equilibrium
    within the FP32 bound.  (The systems alone launch one dependency-chain
    link of two products at a time, as the grouped recurrences do:
    NED_CHAIN=0.)"""
    from nedyson_amd.green import GreenFunc
    monkeypatch.setenv("NED_CHAIN", "0")
    nt = 17
    ps = problem.batch_systems(b_sd=0.05, nt=nt)
    print(f"[batch16] nao {[p.nao for p in ps]}", flush=True)
    alone = []
    eq = []
    for p in ps:
        mu, F, GM = _equilibrium(p)
        eq.append((mu, F, GM))
        c = _equilibrium_ctx(p, nt, mu, F, GM)
        for T in range(p.k + 1, nt + 1):
            c.do_tstp(T, p.h0, mu, p.beta, p.dt, "f32", 3)
        alone.append(c.download(ned.NED_G))
        c.close()
    print("[batch16] alone done", flush=True)
    ctxs = [_equilibrium_ctx(p, nt, *e) for p, e in zip(ps, eq)]
    b = ned.Batch(ctxs)
    for T in range(ps[0].k + 1, nt + 1):
        it, err = b.do_tstp(T, [p.h0 for p in ps], [m for m, _, _ in eq], [p.beta for p in ps],
                            [p.dt for p in ps], "f32", 3)
        assert np.all(np.isfinite(err))
    for p, c, ref in zip(ps, ctxs, alone):
        d = c.download(ned.NED_G)
        for comp in ("ret", "les", "tv"):
            assert np.array_equal(getattr(d, comp), getattr(ref, comp)), \
                (p.nao, comp, np.max(np.abs(getattr(d, comp) - getattr(ref, comp))))
    b.close()
    for c in ctxs:
        c.close()
    i8 = [p.nao for p in ps].index(8)
    p, (mu, F, GM) = ps[i8], eq[i8]
    G = problem.free_gf(F, mu, p.beta, p.dt, nt, p.ntau, p.k, GM)
    S = GreenFunc(nt, p.ntau, p.nao)
    hmf = np.zeros((nt + 1, p.nao, p.nao), dtype=complex)
    D = orc.Dyson(nt, p.ntau, p.nao, p.k)
    D.do_boot(p.V, G, S, hmf, F, mu, p.beta, p.dt, 60, 1e-12)
    for T in range(p.k + 1, nt + 1):
        D.do_tstp(p.V, T, G, S, hmf, p.h0, mu, p.beta, p.dt, 3)
    for comp in ("ret", "les", "tv"):
        err = np.max(np.abs(getattr(alone[i8], comp) - getattr(G, comp)))
        print(f"[batch16] nao=8 vs oracle {comp}: {err:.2e}", flush=True)
        assert err < 1e-6, comp
