"""CPU-side checks of the product library: exported ABI, host quadrature
tables, problem setup.  No GPU compute is called."""
import json
import os
import re

import numpy as np
import pytest

import nedyson_amd as ned
from nedyson_amd import problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "nedyson.h")).read()
    declared = sorted(set(re.findall(r"\b(ned_[a-z_0-9]+)\s*\(", hdr)))
    assert len(declared) >= 20
    L = ned.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(ned.EXPORTS) == declared


def test_product_integrator_against_golden_and_oracle(orc):
    cc = json.load(open(os.path.join(GOLD, "calccoeff.json")))
    P = ned.integrator(5)
    O = orc.integrator(5)
    for key in ("poly_interp", "poly_diff", "poly_integ", "bd", "start", "Omega", "ex"):
        assert np.max(np.abs(P[key] - O[key])) < 1e-14, key
    assert np.linalg.norm(np.array(cc["PDiff"]) - P["poly_diff"]) < 1e-12
    assert np.linalg.norm(np.array(cc["PInteg"]) - P["poly_integ"]) < 1e-12
    assert np.linalg.norm(np.array(cc["BD"]) - P["bd"]) < 1e-12
    assert np.linalg.norm(np.array(cc["w"]) - P["Omega"][30:36]) < 1e-12


@pytest.mark.parametrize("ntau", [5, 30, 63, 127])
def test_gll_nodes_agree(orc, ntau):
    x_o, _ = orc.gll(ntau)
    assert np.max(np.abs(ned.gll_nodes(ntau) - x_o)) < 1e-15
    assert np.max(np.abs(problem.gll_nodes_np(ntau) - x_o)) < 1e-14


def test_hf_equilibrium_is_self_consistent():
    p = problem.Problem(nao=4, naux=12, nelec=2, beta=10.0, ntau=15, nt=10, dt=0.02, seed=0)
    F = problem.hf_fock(p.h1, p.B, p.rho_eq)
    assert np.max(np.abs(F - p.fock)) < 1e-8
    assert abs(2 * np.trace(p.rho_eq) - 2) < 1e-10
    # -G^M(beta) is the density (src/greens.h:56-58)
    assert np.max(np.abs(-p.Gmat[-1].real - p.rho_eq)) < 1e-10
    assert np.allclose(p.B, np.transpose(p.B, (0, 2, 1)))


def test_context_fails_loudly_without_device():
    """No CPU fallback: without a HIP device creating a context is an error."""
    try:
        c = ned.Context(2, 3, 6, 5, 5)
    except ned.NedError as e:
        assert e.code in (1, 3)
        return
    c.close()  # a device is present: creation must then succeed


def test_free_gf_matches_oracle(orc):
    """problem.free_gf (the bootstrap's initial guess) == oracle G0_from_h0."""
    p = problem.Problem(nao=3, naux=5, nelec=2, beta=5.0, ntau=12, nt=8, dt=0.02, seed=4)
    g = p.initial_green()
    from nedyson_amd.green import GreenFunc
    G = GreenFunc(p.nt, p.ntau, p.nao)
    D = orc.Dyson(p.nt, p.ntau, p.nao, p.k)
    D.g0_from_h0(G, p.mu, p.fock, p.beta, p.dt)
    D.g0_matsubara(G, p.mu, p.fock, p.beta)
    for c in ("ret", "les", "tv", "mat"):
        assert np.max(np.abs(getattr(G, c) - getattr(g, c))) < 1e-12, c
