"""Graph mode of ned_do_tstp (ned_set_graph): the corrector passes after the
first of each step replayed from one captured HIP graph give the same G, bit
for bit, as the passes launched one by one.  Every step captures a new graph
and replays it twice, with the first pass and the host's reads of the update
norm (eager work and synchronisations) between the replays."""
import numpy as np
import pytest

import nedyson_amd as ned
from nedyson_amd import problem

pytestmark = pytest.mark.gpu


def _run_small(graph, prec, corr_tol):
    p = problem.Problem(nao=4, naux=12, nelec=2, beta=10.0, ntau=15, nt=30, dt=0.02, seed=5, spacing=1.0, b_sd=0.1,
                        quench="h00", quench_norm=0.5)
    its = []
    with ned.Context(p.nao, p.naux, p.nt, p.ntau, p.k) as c:
        c.set_coulomb(p.B)
        c.upload(ned.NED_G, p.initial_green())
        c.set_matsubara(p.beta)
        c.do_boot(p.fock, p.mu, p.beta, p.dt, "f64", 50, 1e-12)
        c.set_graph(graph)
        for T in range(p.k + 1, p.nt + 1):
            its.append(c.do_tstp(T, p.h0, p.mu, p.beta, p.dt, prec, 4 if corr_tol else 3, corr_tol))
        return c.download(ned.NED_G), its


@pytest.mark.parametrize("prec,corr_tol", [("f64", 0.0), ("bf16x3", 0.0), ("f64", 1e-9)])
def test_graph_mode_matches_launched_passes_small(prec, corr_tol):
    g0, it0 = _run_small(False, prec, corr_tol)
    g1, it1 = _run_small(True, prec, corr_tol)
    assert [i for i, _ in it0] == [i for i, _ in it1]
    for comp in ("ret", "les", "tv"):
        assert np.array_equal(getattr(g0, comp), getattr(g1, comp)), comp


@pytest.mark.parametrize("prec", ["f64", "bf16x3"])
def test_graph_mode_matches_launched_passes_nao64(prec, monkeypatch):
    """nao = 64, naux = 64: the batched 64 x 64 Sigma chain with the split
    exchange in bf16x3 and the band-split Dyson recurrences with their side
    streams, on a 40-step free-G history and 3 steps, graph against launched."""
    monkeypatch.setenv("NED_NT_VARIANT", "11")  # one tile variant for both runs
    p = problem.Problem(nao=64, naux=64, nelec=32, beta=4.0, ntau=15, nt=43, dt=0.01, seed=64, spacing=0.1,
                        b_sd=0.02, quench="random", quench_norm=0.2)
    T0 = 40
    outs = []
    for graph in (False, True):
        print(f"[graph] nao=64 {prec} graph={graph}", flush=True)
        with ned.Context(p.nao, p.naux, p.nt, p.ntau, p.k) as c:
            c.set_coulomb(p.B)
            c.write_blocks(ned.NED_G, ned.NED_MAT, 0, p.Gmat)
            c.set_matsubara(p.beta)
            c.g0_from_h0(T0, p.fock, p.mu, p.beta, p.dt)
            c.write_hmf(np.repeat(p.fock[None].astype(complex), T0 + 1, axis=0))
            c.sigma(T0, "f64")
            c.tti_extend(ned.NED_SIGMA, T0)
            c.set_graph(graph)
            for T in range(T0 + 1, p.nt + 1):
                c.do_tstp(T, p.h0, p.mu, p.beta, p.dt, prec, 3, 0.0)
            outs.append(c.download(ned.NED_G))
    for comp in ("ret", "les", "tv"):
        assert np.array_equal(getattr(outs[0], comp), getattr(outs[1], comp)), comp
