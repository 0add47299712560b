"""Loaders for the reference's Dyson-solver and free-GF test data, as
exported to tests/golden/reference_dyson_fixtures.npz by make_golden.py.

dyson_test_data.h5 holds a nao = 2 system with nt = 20, ntau = 30: the
self-energy it was propagated with (Sigma), the propagated G (Timesteps), an
extrapolation run (Extrap) and the mean field (epsmf/ft, 22 entries).  The
entries match the solver with mu = 1, beta = 9, dt = 0.1, k = 5 and
hmf[t] = ft[t + 1]; its imaginary-time data are sampled on a uniform tau
grid (an earlier convention of the reference), so only the parts that do
not involve tau integrals -- the retarded rows and the extrapolation -- are
comparable with the GLL-grid solver.  Gfree_const.h5 is the golden file of
tests/free-test.cpp:8-44 (G0_from_h0 with mu = 1, beta = 9, dt = 0.1,
h0 = [[1, -3], [-3, 2]])."""
import os

import numpy as np

from nedyson_amd.green import GreenFunc

NPZ = os.path.join(os.path.dirname(__file__), "golden", "reference_dyson_fixtures.npz")
NT, NTAU, NAO, K = 20, 30, 2, 5
MU, BETA, DT = 1.0, 9.0, 0.1
H0_FREE = np.array([[1.0, -3.0], [-3.0, 2.0]])


def _data():
    return np.load(NPZ, allow_pickle=False)


def green(group):
    """GreenFunc of group Timesteps, Sigma, Extrap or Gfree."""
    d = _data()
    g = GreenFunc(NT, NTAU, NAO)
    g.ret[:] = d[f"{group}_GR"].reshape(g.ret.shape)
    g.les[:] = d[f"{group}_GL"].reshape(g.les.shape)
    g.tv[:] = d[f"{group}_GTV"].reshape(g.tv.shape)
    g.mat[:] = d[f"{group}_GM"].reshape(g.mat.shape)
    return g


def hmf():
    """The mean field of the Timesteps run, hmf[t] = ft[t + 1], t = 0..NT."""
    ft = _data()["epsmf_ft"].reshape(NT + 2, NAO, NAO)
    return np.ascontiguousarray(ft[1:NT + 2])
