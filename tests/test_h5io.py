"""The GREEN HDF5 reader/writer (libnedyson_h5.so) on the host: the
reference's own GREEN file (tests/data/Gfree_const.h5, written by
green_func::print_to_file), round trips, and the layout as h5py sees it."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from nedyson_amd.green import GreenFunc

GOLD = os.path.join(os.path.dirname(__file__), "golden")
h5 = pytest.importorskip("nedyson_amd.h5")


def _lib_or_skip():
    try:
        h5.lib()
    except ImportError as e:  # HDF5 headers absent at build time
        pytest.skip(str(e))


def test_reads_reference_green_file():
    """Sizes and every block of the reference's Gfree_const.h5 (compared with
    the copy exported through h5py into reference_dyson_fixtures.npz)."""
    _lib_or_skip()
    import reference_fixtures as RF
    path = os.path.join(GOLD, "Gfree_const.h5")
    assert h5.dims(path) == (RF.NT, RF.NTAU, RF.NAO, -1)
    g = h5.read_green(path)
    ref = RF.green("Gfree")
    for a, b in ((g.ret, ref.ret), (g.les, ref.les), (g.tv, ref.tv), (g.mat, ref.mat)):
        assert np.array_equal(a, b)


def test_round_trip_groups_and_overwrite(tmp_path):
    _lib_or_skip()
    rng = np.random.default_rng(4)
    path = str(tmp_path / "g.h5")

    def rand(nt, ntau, nao):
        g = GreenFunc(nt, ntau, nao)
        for a in (g.ret, g.les, g.tv, g.mat):
            a[:] = rng.normal(size=a.shape) + 1j * rng.normal(size=a.shape)
        return g
    a, b, c = rand(7, 5, 3), rand(4, 9, 2), rand(7, 5, 3)
    h5.write_green(a, path, "G")
    h5.write_green(b, path, "run/Sigma")
    assert h5.dims(path, "G") == (7, 5, 3, -1) and h5.dims(path, "run/Sigma") == (4, 9, 2, -1)
    for g, grp in ((a, "G"), (b, "run/Sigma")):
        r = h5.read_green(path, grp)
        for x, y in ((r.ret, g.ret), (r.les, g.les), (r.tv, g.tv), (r.mat, g.mat)):
            assert np.array_equal(x, y)
    h5.write_green(c, path, "G")  # the group is replaced in place
    r = h5.read_green(path, "G")
    assert np.array_equal(r.tv, c.tv) and np.array_equal(h5.read_green(path, "run/Sigma").mat, b.mat)
    with pytest.raises(h5.H5Error):
        h5.read_green(path, "nothing")


def test_layout_matches_reference_format(tmp_path):
    """Datasets GR, GL, GTV, GM as flat complex128 of the packed sizes and int
    scalars nt, ntau, sig, nao, as h5py reads them (the interpreter with h5py
    in this image, when present)."""
    _lib_or_skip()
    py = "/opt/conda/bin/python3.9"
    if not os.path.exists(py):
        pytest.skip("no interpreter with h5py")
    g = GreenFunc(6, 4, 2)
    g.ret[:] = 1 + 2j
    path = str(tmp_path / "x.h5")
    h5.write_green(g, path)
    code = ("import h5py,sys; f=h5py.File(sys.argv[1],'r'); "
            "print(sorted(f.keys())); "
            "print([(k, f[k].shape, str(f[k].dtype)) for k in ('GR','GL','GTV','GM')]); "
            "print([int(f[k][()]) for k in ('nt','ntau','sig','nao')]); print(f['GR'][0])")
    r = subprocess.run([py, "-c", code, path], capture_output=True, text=True, timeout=60)
    if r.returncode != 0:
        pytest.skip("h5py unavailable: " + r.stderr[-200:])
    lines = r.stdout.strip().splitlines()
    assert lines[0] == str(sorted(["GL", "GM", "GR", "GTV", "nao", "nt", "ntau", "sig"]))
    assert lines[1] == str([("GR", (28 * 4,), "complex128"), ("GL", (28 * 4,), "complex128"),
                            ("GTV", (7 * 5 * 4,), "complex128"), ("GM", (5 * 4,), "complex128")])
    assert lines[2] == "[6, 4, -1, 2]"
    assert lines[3] == "(1+2j)"


@pytest.mark.gpu
def test_context_round_trip_and_reference_file(tmp_path):
    """Device contexts: G and Sigma written by one context and read by
    another; a partial write (slices 0..3); the reference's file read into a
    context."""
    _lib_or_skip()
    import nedyson_amd as ned
    import reference_fixtures as RF
    rng = np.random.default_rng(5)
    nt, ntau, nao = 6, 5, 3
    gs = []
    for _ in range(2):
        g = GreenFunc(nt, ntau, nao)
        for a in (g.ret, g.les, g.tv, g.mat):
            a[:] = rng.normal(size=a.shape) + 1j * rng.normal(size=a.shape)
        gs.append(g)
    path = str(tmp_path / "ctx.h5")
    with ned.Context(nao, 2, nt, ntau, 5) as a, ned.Context(nao, 2, nt, ntau, 5) as b:
        a.upload(ned.NED_G, gs[0])
        a.upload(ned.NED_SIGMA, gs[1])
        h5.write_context(a, path, "G", ned.NED_G)
        h5.write_context(a, path, "Sigma", ned.NED_SIGMA)
        h5.write_context(a, path, "G3", ned.NED_G, nt=3)
        h5.read_context(b, path, "G", ned.NED_G)
        h5.read_context(b, path, "Sigma", ned.NED_SIGMA)
        for which, g in ((ned.NED_G, gs[0]), (ned.NED_SIGMA, gs[1])):
            r = b.download(which)
            for x, y in ((r.ret, g.ret), (r.les, g.les), (r.tv, g.tv), (r.mat, g.mat)):
                assert np.array_equal(x, y)
    assert h5.dims(path, "G3") == (3, ntau, nao, -1)
    g3 = h5.read_green(path, "G3")
    assert np.array_equal(g3.ret, gs[0].ret[:10]) and np.array_equal(g3.tv, gs[0].tv[:4])
    with ned.Context(RF.NAO, 1, RF.NT, RF.NTAU, RF.K) as c:
        h5.read_context(c, os.path.join(GOLD, "Gfree_const.h5"))
        r = c.download(ned.NED_G)
    ref = RF.green("Gfree")
    assert np.array_equal(r.ret, ref.ret) and np.array_equal(r.mat, ref.mat)


def test_h5_library_exports_every_declared_symbol():
    import re
    _lib_or_skip()
    hdr = open(os.path.join(os.path.dirname(GOLD), "..", "include", "nedyson_h5.h")).read()
    declared = sorted(set(re.findall(r"\b(ned_h5_[a-z_0-9]+)\s*\(", hdr)))
    assert len(declared) == 6
    for name in declared:
        assert hasattr(h5.lib(), name), name
