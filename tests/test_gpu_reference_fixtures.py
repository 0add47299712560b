"""The GPU path against the reference's own stored outputs
(tests/golden/reference_dyson_fixtures.npz): dyson_step's retarded rows and
Extrapolate from tests/data/dyson_test_data.h5, and G0_from_h0 from
tests/data/Gfree_const.h5 (tests/free-test.cpp:8-44)."""
import numpy as np
import pytest

import nedyson_amd as ned
import reference_fixtures as RF
from nedyson_amd.problem import fermi

pytestmark = pytest.mark.gpu


def _context():
    c = ned.Context(RF.NAO, 1, RF.NT, RF.NTAU, RF.K)
    c.set_coulomb(np.zeros((1, RF.NAO, RF.NAO)))
    return c


def test_gpu_dyson_step_ret_against_reference_timesteps():
    """ned_dyson_step from the stored history, T = k+1..20: the retarded row
    within 1e-12 of the stored one (Sigma != 0, time-dependent mean field)."""
    S, Gt = RF.green("Sigma"), RF.green("Timesteps")
    with _context() as c:
        c.upload(ned.NED_SIGMA, S)
        c.write_hmf(RF.hmf().astype(np.complex128))
        worst = 0.0
        for T in range(RF.K + 1, RF.NT + 1):
            c.upload(ned.NED_G, Gt)  # the stored history (writing G^M drops the mixing kernel)
            c.set_matsubara(RF.BETA)
            c.dyson_step(T, RF.MU, RF.BETA, RF.DT)
            row = c.read_blocks(ned.NED_G, ned.NED_RET, T * (T + 1) // 2, T + 1)
            err = np.max(np.abs(row - Gt.ret_row(T)))
            worst = max(worst, err)
            assert err < 1e-12, (T, err)
    print(f"[reference] dyson_step retarded rows: max |dG| {worst:.2e}", flush=True)


def test_gpu_extrapolate_against_reference():
    """ned_extrapolate: every slice of the stored extrapolation run from the
    slices before it (all components, 1e-12 relative)."""
    Ge = RF.green("Extrap")
    with _context() as c:
        for T in range(RF.K + 1, RF.NT + 1):
            c.upload(ned.NED_G, Ge)
            c.extrapolate(T)
            s = T * (T + 1) // 2
            got = (c.read_blocks(ned.NED_G, ned.NED_RET, s, T + 1), c.read_blocks(ned.NED_G, ned.NED_LES, s, T + 1),
                   c.read_blocks(ned.NED_G, ned.NED_TV, T * (RF.NTAU + 1), RF.NTAU + 1))
            for a, b in zip(got, (Ge.ret_row(T), Ge.les_col(T), Ge.tv[T])):
                assert np.max(np.abs(a - b)) <= 1e-12 * max(1.0, np.max(np.abs(b))), T


def test_gpu_g0_from_h0_against_reference_free_test():
    """ned_g0_from_h0 (k steps) against Gfree_const.h5: G^R and G^< directly,
    G^tv against the closed form that reproduces the file's own (uniform-grid)
    G^tv, here on the GLL nodes."""
    ref = RF.green("Gfree")
    nb = (RF.K + 1) * (RF.K + 2) // 2
    with _context() as c:
        c.g0_from_h0(RF.K, RF.H0_FREE, RF.MU, RF.BETA, RF.DT)
        ret = c.read_blocks(ned.NED_G, ned.NED_RET, 0, nb)
        les = c.read_blocks(ned.NED_G, ned.NED_LES, 0, nb)
        tv = c.read_blocks(ned.NED_G, ned.NED_TV, 0, (RF.K + 1) * (RF.NTAU + 1)).reshape(RF.K + 1, RF.NTAU + 1, 2, 2)
    assert np.max(np.abs(ret - ref.ret[:nb])) < 1e-12
    assert np.max(np.abs(les - ref.les[:nb])) < 1e-12
    lam, V = np.linalg.eigh(RF.MU * np.eye(RF.NAO) - RF.H0_FREE)
    x = ned.gll_nodes(RF.NTAU)

    def fexp(t, w):
        return np.where(w < 0, np.exp(w * t) * fermi(RF.BETA, w), np.exp((t - RF.BETA) * w) * fermi(RF.BETA, -w))
    for m in range(RF.K + 1):
        U = V @ np.diag(np.exp(1j * RF.DT * m * lam)) @ V.T
        cf = np.array([1j * U @ ((V * fexp(t, -lam)) @ V.T) for t in RF.BETA * (np.asarray(x) + 1) / 2])
        assert np.max(np.abs(tv[m] - cf)) < 1e-12
