"""Write the reference's own golden data for this path as small JSON fixtures.

Run with an interpreter that has h5py (e.g. /opt/conda/bin/python3.9) in a
container where the reference checkout is mounted at /root/reference:

    /opt/conda/bin/python3.9 tests/golden/make_golden.py

Outputs (committed):
  calccoeff.json   -- tests/data/calccoeff.h5 (Integrator(5) tables pinned by
                      tests/integration-test.cpp:5-69)
  integrals.json   -- the expected scalars of tests/integrals-test.cpp:8-212
                      (energy_conv, CTV1, Ctv_tstp, Cles2_tstp, Cles3_tstp on
                      the closed-form contour functions of tests/set_up_g.h)
  dsim_invariants.json -- slices of tests/data/Dsim.h5 (a decomposed-GF2
                      equilibrium run, nao=2, nt=ntau=100): G^M (on a uniform
                      tau grid), the first retarded column, the lesser
                      diagonal and a few more blocks.
  h2_hf.json       -- tests/data/hf.hdf5 (h1, Fock matrix, Coulomb integrals
                      (ij|kl) of the H2 system Dsim.h5 propagates).
  reference_dyson_fixtures.npz -- the contour functions of
                      tests/data/dyson_test_data.h5 (Timesteps, Sigma, Extrap,
                      epsmf/ft) and tests/data/Gfree_const.h5 (free-test.cpp).
  Gfree_const.h5   -- tests/data/Gfree_const.h5 itself (a GREEN file written by
                      the reference), read by the HDF5 GREEN reader's tests.
"""
import json
import os
import sys

import numpy as np

REF = "/root/reference/tests/data"
OUT = os.path.dirname(os.path.abspath(__file__))


def calccoeff():
    import h5py
    f = h5py.File(os.path.join(REF, "calccoeff.h5"), "r")
    d = {k: np.asarray(f[k][()], dtype=float).tolist() for k in f}
    d["_source"] = "tests/data/calccoeff.h5 (checked by tests/integration-test.cpp:5-69)"
    return d


def integrals():
    # Values transcribed from tests/integrals-test.cpp (line numbers given).
    # Each entry: [argument, real, imag, tolerance]
    return {
        "_source": "tests/integrals-test.cpp:8-212; inputs tests/set_up_g.h:13-243",
        "params": {"Nt": 50, "Ntau": 30, "Nao": 2, "beta": 1.5, "dt": 0.02, "k": 5},
        "energy_conv": [  # :21-41, set_up_G_herm, returns a real number
            [0, -0.28125, 0.0, 1e-11],
            [3, -0.2927488689451264, 0.0, 1e-10],
            [5, -0.30327566284415913, 0.0, 1e-10],
            [7, -0.315908825747133, 0.0, 1e-10],
            [8, -0.3229469844098787, 0.0, 1e-10],
            [10, -0.33829858325203477, 0.0, 1e-10],
            [20, -0.42913160446733745, 0.0, 1e-10],
        ],
        "CTV1": [  # :61-85, sum over all entries of CTV1(res, S, Scc, G, n, dt)
            [0, 0.0, 0.0, 1e-11],
            [2, 1.942981791711302, 0.1785128961274855, 1e-11],
            [4, 3.8094748281637, 0.7143691514656275, 1e-11],
            [7, 6.472103613463175, 2.190432440083725, 1e-11],
            [9, 8.16023388825266, 3.62521550316101, 1e-11],
            [20, 16.37804189025634, 18.11528562371781, 1e-11],
        ],
        "Ctv_tstp": [  # :100-124, sum of G.tv(tstp, :) after Ctv_tstp(tstp, G, G, Gcc, S, Scc)
            [0, 59.99682203389833, 0.0, 1e-11],
            [2, 59.11788717375197, -0.06479860320481823, 1e-11],
            [4, 58.64718604482833, 0.03667453794393438, 1e-10],
            [7, 58.78029162902846, 0.5003153197597525, 1e-10],
            [9, 59.47542129291611, 1.016421704133382, 1e-10],
            [20, 73.36535591913849, 6.770356696009172, 1e-10],
        ],
        "Cles2_tstp": [  # :143-167, sum of the first max(m,5)+1 blocks of res
            [0, 0.0, 0.0, 1e-11],
            [2, -0.9779017217310684, -0.004556524033142945, 1e-11],
            [5, -2.506405029994643, -0.028363125672494226, 1e-11],
            [6, -3.534156697360735, -0.04705129775222273, 1e-11],
            [9, -7.715189412549618, -0.14521215747129892, 1e-11],
            [20, -37.224605902254595, -1.258015308782273, 1e-11],
        ],
        "Cles3_tstp": [  # :186-210, sum of the first max(m,5)+1 blocks of res
            [0, 6.427012997367941, 5.348335104615002, 1e-11],
            [2, 6.64094640155254, 5.091254584720283, 1e-11],
            [5, 6.961846507829439, 4.705633804878206, 1e-11],
            [6, 8.42467270546969, 5.318615593818669, 1e-11],
            [9, 13.32282738165199, 6.804593007976855, 1e-11],
            [20, 37.65576268035543, 6.65193545259473, 1e-11],
        ],
    }


def dsim():
    import h5py
    f = h5py.File(os.path.join(REF, "Dsim.h5"), "r")
    g = f["G"]
    nt, ntau, nao = int(g["nt"][()]), int(g["ntau"][()]), int(g["nao"][()])
    es = nao * nao
    GL = g["GL"][()].reshape(-1, nao, nao)
    GR = g["GR"][()].reshape(-1, nao, nao)
    GTV = g["GTV"][()].reshape(nt + 1, ntau + 1, nao, nao)
    GM = g["GM"][()].reshape(ntau + 1, nao, nao)

    def les(i, j):
        return GL[j * (j + 1) // 2 + i]

    def ret(i, j):
        return GR[i * (i + 1) // 2 + j]

    out = {"_source": "tests/data/Dsim.h5 group G", "nt": nt, "ntau": ntau, "nao": nao, "sig": int(g["sig"][()])}
    # a small slice of the stored data, so parity with the packing is testable
    out["GM"] = [[c.real, c.imag] for c in GM.reshape(-1)]
    out["GTV_t0"] = [[c.real, c.imag] for c in GTV[0].reshape(-1)]
    out["GTV_t5"] = [[c.real, c.imag] for c in GTV[5].reshape(-1)]
    out["GR_diag"] = [[c.real, c.imag] for t in range(0, nt + 1, 10) for c in ret(t, t).reshape(-1)]
    out["GL_5_10"] = [[c.real, c.imag] for c in les(5, 10).reshape(-1)]
    out["GR_10_5"] = [[c.real, c.imag] for c in ret(10, 5).reshape(-1)]
    out["density"] = [float(np.real(np.trace(1j * -1 * les(t, t)))) * 2 for t in range(0, nt + 1, 10)]
    # every time step of the first retarded column and of the lesser diagonal
    out["GR_t0"] = [[c.real, c.imag] for t in range(nt + 1) for c in ret(t, 0).reshape(-1)]
    out["GL_tt"] = [[c.real, c.imag] for t in range(nt + 1) for c in les(t, t).reshape(-1)]
    return out


def h2_system():
    """tests/data/hf.hdf5: the restricted-HF H2 system (minimal basis) whose
    equilibrium and propagation Dsim.h5 holds."""
    import h5py
    f = h5py.File(os.path.join(REF, "hf.hdf5"), "r")
    out = {"_source": "tests/data/hf.hdf5"}
    for k in ("hcore", "fock", "uchem"):
        out[k] = np.asarray(f[k][()], dtype=float).reshape(-1).tolist()
    out["nelectron"] = int(f["nelectron"][()])
    out["enuc"] = float(f["enuc"][()])
    out["nao"] = int(f["hcore"].shape[0])
    return out


def dyson_fixtures():
    """Contour functions of the reference's Dyson-solver test data
    (tests/data/dyson_test_data.h5: groups Timesteps, Sigma, Extrap and the
    mean field epsmf/ft; nao = 2, nt = 20, ntau = 30) and of the free-GF test
    (tests/data/Gfree_const.h5, checked by tests/free-test.cpp:8-44)."""
    import h5py
    out = {}
    f = h5py.File(os.path.join(REF, "dyson_test_data.h5"), "r")
    for g in ("Timesteps", "Sigma", "Extrap"):
        for c in ("GR", "GL", "GTV", "GM"):
            out[f"{g}_{c}"] = np.asarray(f[g][c][()], dtype=np.complex128)
    out["epsmf_ft"] = np.asarray(f["epsmf"]["ft"][()], dtype=np.complex128)
    f = h5py.File(os.path.join(REF, "Gfree_const.h5"), "r")
    for c in ("GR", "GL", "GTV", "GM"):
        out[f"Gfree_{c}"] = np.asarray(f[c][()], dtype=np.complex128)
    return out


def main():
    json.dump(calccoeff(), open(os.path.join(OUT, "calccoeff.json"), "w"), indent=1)
    json.dump(integrals(), open(os.path.join(OUT, "integrals.json"), "w"), indent=1)
    try:
        json.dump(dsim(), open(os.path.join(OUT, "dsim_invariants.json"), "w"), indent=1)
    except Exception as e:  # pragma: no cover
        print("Dsim.h5 not read:", e, file=sys.stderr)
    json.dump(h2_system(), open(os.path.join(OUT, "h2_hf.json"), "w"), indent=1)
    np.savez_compressed(os.path.join(OUT, "reference_dyson_fixtures.npz"), **dyson_fixtures())
    # one of the reference's GREEN files as it is, for the HDF5 reader
    import shutil
    shutil.copyfile(os.path.join(REF, "Gfree_const.h5"), os.path.join(OUT, "Gfree_const.h5"))


if __name__ == "__main__":
    main()
