"""GPU parity of the second-Born self-energy (ned_sigma_gf2_decomp) with the
CPU oracle's restatement of molGF2SolverDecomp::solve."""
import numpy as np
import pytest

import nedyson_amd as ned
from helpers import random_green, sym_V
from nedyson_amd.green import GreenFunc

pytestmark = pytest.mark.gpu

# max-abs error relative to max |Sigma| (FP64 is checked absolutely below)
TOL = {"f64": 1e-12, "f32": 2e-5, "bf16x3": 2e-4}


def _sigma_gpu(V, B, G, tstp, prec, chunk=None):
    nao, naux = V.shape[0], V.shape[2]
    with ned.Context(nao, naux, G.nt, G.ntau, 1) as c:
        c.set_coulomb(B)
        c.upload(ned.NED_G, G)
        c.sigma(tstp, prec)
        return c.download(ned.NED_SIGMA)


# (16, 64) and (32, 64) take the fused split-bf16 path in bf16x3 (nao^2 and
# 2 naux multiples of 128): Pi on the NT kernel, Pi V folded into the exchange
@pytest.mark.parametrize("nao,naux", [(2, 3), (4, 12), (8, 32), (13, 20), (16, 64), (32, 64)])
@pytest.mark.parametrize("prec", ["f64", "f32", "bf16x3"])
def test_sigma_parity(orc, nao, naux, prec):
    print(f"[sigma parity] nao={nao} naux={naux} prec={prec}", flush=True)
    rng = np.random.default_rng(100 * nao + naux)
    V, B = sym_V(rng, nao, naux)
    nt, ntau = 6, 7
    G = random_green(rng, nt, ntau, nao, scale=0.5)
    for tstp in (0, 3, nt):
        S_ref = GreenFunc(nt, ntau, nao)
        orc.sigma_solve(V, tstp, S_ref, G)
        S = _sigma_gpu(V, B, G, tstp, prec)
        s = tstp * (tstp + 1) // 2
        for name, a, b in (("ret", S.ret_row(tstp), S_ref.ret_row(tstp)),
                           ("les", S.les_col(tstp), S_ref.les_col(tstp)),
                           ("tv", S.tv[tstp], S_ref.tv[tstp])):
            scale = max(np.max(np.abs(b)), 1e-30)
            err = np.max(np.abs(a - b))
            assert err <= TOL[prec] * scale, (name, tstp, err, scale)
        del s


@pytest.mark.parametrize("nao,naux,prec", [(6, 10, "f64"), (16, 64, "bf16x3")])
def test_sigma_chunking_invariant(orc, nao, naux, prec):
    """Splitting the slice into item chunks must not change the result."""
    rng = np.random.default_rng(5)
    nt, ntau = 9, 11
    V, B = sym_V(rng, nao, naux)
    G = random_green(rng, nt, ntau, nao, scale=0.5)
    outs = []
    import os
    for chunk in (0, 1, 5):
        os.environ["NED_SIGMA_CHUNK"] = str(chunk)
        try:
            with ned.Context(nao, naux, nt, ntau, 1) as c:
                c.set_coulomb(B)
                c.upload(ned.NED_G, G)
                c.sigma(nt, prec)
                outs.append(c.download(ned.NED_SIGMA))
        finally:
            del os.environ["NED_SIGMA_CHUNK"]
    for o in outs[1:]:
        for comp in ("ret", "les", "tv"):
            assert np.max(np.abs(getattr(o, comp) - getattr(outs[0], comp))) == 0


def test_sigma_rejects_nonsymmetric_V():
    rng = np.random.default_rng(1)
    B = rng.normal(size=(3, 2, 2))
    with ned.Context(2, 3, 6, 5, 1) as c:
        with pytest.raises(ned.NedError):
            c.set_coulomb(B)


def test_sigma_wrapper_reference_signature(orc):
    """molGF2SolverDecomp(Vija, Viaj).solve(tstp, Sigma, G) on host objects."""
    rng = np.random.default_rng(2)
    nao, naux, nt, ntau = 3, 5, 5, 6
    V, B = sym_V(rng, nao, naux)
    G = random_green(rng, nt, ntau, nao, scale=0.5)
    solver = ned.molGF2SolverDecomp(V)
    S, S_ref = GreenFunc(nt, ntau, nao), GreenFunc(nt, ntau, nao)
    for t in range(nt + 1):
        solver.solve(t, S, G)
        orc.sigma_solve(V, t, S_ref, G)
    for c in ("ret", "les", "tv"):
        assert np.max(np.abs(getattr(S, c) - getattr(S_ref, c))) < 1e-12


def test_sigma_parity_nao64(orc):
    """nao = 64: the bf16x3 chain runs on the batched 64 x 64 kernels (W, Z,
    C, E) with the fused exchange GEMM; FP32 and FP64 take the generic path.
    One slice against the oracle (OpenMP over its time points)."""
    import os
    rng = np.random.default_rng(64)
    nao, naux, nt, ntau = 64, 64, 2, 3
    V, B = sym_V(rng, nao, naux)
    G = random_green(rng, nt, ntau, nao, scale=0.2)
    S_ref = GreenFunc(nt, ntau, nao)
    orc.sigma_solve(V, nt, S_ref, G, nthreads=min(os.cpu_count() or 1, 16))
    for prec in ("bf16x3", "f32", "f64"):
        print(f"[sigma parity] nao={nao} naux={naux} prec={prec}", flush=True)
        S = _sigma_gpu(V, B, G, nt, prec)
        for name, a, b in (("ret", S.ret_row(nt), S_ref.ret_row(nt)),
                           ("les", S.les_col(nt), S_ref.les_col(nt)),
                           ("tv", S.tv[nt], S_ref.tv[nt])):
            scale = max(np.max(np.abs(b)), 1e-30)
            err = np.max(np.abs(a - b))
            assert err <= TOL[prec] * scale, (prec, name, err, scale)


@pytest.mark.parametrize("symx", ["1", "0"])
def test_sigma_every_gemm_variant_nao64(orc, monkeypatch, symx):
    """Every tile variant of the split-bf16 exchange and Pi GEMMs, pinned in
    turn with NED_NT_VARIANT, on the nao = 64 slice of
    test_sigma_parity_nao64 (M = 1280): each against the oracle within the
    bf16x3 bound.  With NED_SYMX=1 (the default) the exchange runs as its
    symmetric and antisymmetric parts under (m,q) -> (q,m) (N = 2304 and
    2048); with NED_SYMX=0 as the full contraction (N = 4096)."""
    import os
    monkeypatch.setenv("NED_SYMX", symx)
    names = ned.gemm_variants()
    assert len(names) >= 10
    rng = np.random.default_rng(64)
    nao, naux, nt, ntau = 64, 64, 2, 3
    V, B = sym_V(rng, nao, naux)
    G = random_green(rng, nt, ntau, nao, scale=0.2)
    S_ref = GreenFunc(nt, ntau, nao)
    orc.sigma_solve(V, nt, S_ref, G, nthreads=min(os.cpu_count() or 1, 16))
    for v, name in enumerate(names):
        print(f"[gemm variant] {v}: {name}", flush=True)
        monkeypatch.setenv("NED_NT_VARIANT", str(v))
        S = _sigma_gpu(V, B, G, nt, "bf16x3")
        for comp, a, b in (("ret", S.ret_row(nt), S_ref.ret_row(nt)),
                           ("les", S.les_col(nt), S_ref.les_col(nt)),
                           ("tv", S.tv[nt], S_ref.tv[nt])):
            scale = max(np.max(np.abs(b)), 1e-30)
            err = np.max(np.abs(a - b))
            assert err <= TOL["bf16x3"] * scale, (v, name, comp, err, scale)


def test_sigma_parity_nao64_naux256(orc):
    """The benchmark's shape, nao = 64 and naux = 256: the split exchange with
    K = 2080 + 256 padded to 2368 and the Pi block in its symmetric operand, one
    slice against the oracle (OpenMP over its time points) in bf16x3."""
    import os
    rng = np.random.default_rng(256)
    nao, naux, nt, ntau = 64, 256, 2, 3
    V, B = sym_V(rng, nao, naux)
    G = random_green(rng, nt, ntau, nao, scale=0.2)
    S_ref = GreenFunc(nt, ntau, nao)
    orc.sigma_solve(V, nt, S_ref, G, nthreads=min(os.cpu_count() or 1, 16))
    print(f"[sigma parity] nao={nao} naux={naux} prec=bf16x3", flush=True)
    S = _sigma_gpu(V, B, G, nt, "bf16x3")
    for name, a, b in (("ret", S.ret_row(nt), S_ref.ret_row(nt)),
                       ("les", S.les_col(nt), S_ref.les_col(nt)),
                       ("tv", S.tv[nt], S_ref.tv[nt])):
        scale = max(np.max(np.abs(b)), 1e-30)
        err = np.max(np.abs(a - b))
        assert err <= TOL["bf16x3"] * scale, (name, err, scale)


@pytest.mark.parametrize("naux", [64, 256])
def test_sigma_chunking_invariant_nao64(orc, monkeypatch, naux):
    """The split exchange at nao = 64 under item chunks of 3 and 7 as well as
    the whole slice (10 items): chunks starting at p0 > 0, a smaller last
    chunk, the W planes kept in H across a chunk, and the Sigma^R of a ">"
    item whose "<" partner fell into an earlier chunk.  With one tile variant
    pinned the result is bitwise the same for every chunking, and within the
    bf16x3 bound of the oracle."""
    import os
    monkeypatch.setenv("NED_SYMX", "1")
    monkeypatch.setenv("NED_NT_VARIANT", "11")  # 128 x 128: fits every chunk's M = 2 naux P
    rng = np.random.default_rng(640 + naux)
    nao, nt, ntau = 64, 2, 3
    V, B = sym_V(rng, nao, naux)
    G = random_green(rng, nt, ntau, nao, scale=0.2)
    S_ref = GreenFunc(nt, ntau, nao)
    orc.sigma_solve(V, nt, S_ref, G, nthreads=min(os.cpu_count() or 1, 16))
    outs = []
    for chunk in (0, 3, 7):
        print(f"[sigma chunking] nao=64 naux={naux} chunk={chunk}", flush=True)
        monkeypatch.setenv("NED_SIGMA_CHUNK", str(chunk))
        outs.append(_sigma_gpu(V, B, G, nt, "bf16x3"))
    for o in outs[1:]:
        for comp in ("ret", "les", "tv"):
            assert np.array_equal(getattr(o, comp), getattr(outs[0], comp)), comp
    S = outs[0]
    for name, a, b in (("ret", S.ret_row(nt), S_ref.ret_row(nt)),
                       ("les", S.les_col(nt), S_ref.les_col(nt)),
                       ("tv", S.tv[nt], S_ref.tv[nt])):
        scale = max(np.max(np.abs(b)), 1e-30)
        assert np.max(np.abs(a - b)) <= TOL["bf16x3"] * scale, name


def test_sigma_parity_nao64_naux256_f32(orc):
    """FP32 Sigma at the benchmark's shape (nao = 64, naux = 256), one slice
    against the oracle."""
    import os
    rng = np.random.default_rng(257)
    nao, naux, nt, ntau = 64, 256, 2, 3
    V, B = sym_V(rng, nao, naux)
    G = random_green(rng, nt, ntau, nao, scale=0.2)
    S_ref = GreenFunc(nt, ntau, nao)
    orc.sigma_solve(V, nt, S_ref, G, nthreads=min(os.cpu_count() or 1, 16))
    print(f"[sigma parity] nao={nao} naux={naux} prec=f32", flush=True)
    S = _sigma_gpu(V, B, G, nt, "f32")
    for name, a, b in (("ret", S.ret_row(nt), S_ref.ret_row(nt)),
                       ("les", S.les_col(nt), S_ref.les_col(nt)),
                       ("tv", S.tv[nt], S_ref.tv[nt])):
        scale = max(np.max(np.abs(b)), 1e-30)
        err = np.max(np.abs(a - b))
        assert err <= TOL["f32"] * scale, (name, err, scale)


def test_nt_autotune_ragged_rows(orc):
    """The split-bf16 GEMM's tile choice is cached per shape class: a row count
    that only 128-row tiles divide (M = 2 naux * 17 = 2176) must not reuse a
    256-row choice made for one in the same octave (M = 2048).  nao = 16,
    naux = 64 (the fused split-bf16 path), slices of 16 and 17 items (ntau+1 =
    14 and 15 at tstp = 0) in one process, each against the oracle."""
    rng = np.random.default_rng(17)
    nao, naux, nt = 16, 64, 6
    V, B = sym_V(rng, nao, naux)
    for ntau in (13, 14):
        G = random_green(rng, nt, ntau, nao, scale=0.5)
        S_ref = GreenFunc(nt, ntau, nao)
        orc.sigma_solve(V, 0, S_ref, G)
        print(f"[autotune rows] items={2 + ntau + 1}", flush=True)
        S = _sigma_gpu(V, B, G, 0, "bf16x3")
        for name, a, b in (("ret", S.ret_row(0), S_ref.ret_row(0)),
                           ("les", S.les_col(0), S_ref.les_col(0)),
                           ("tv", S.tv[0], S_ref.tv[0])):
            scale = max(np.max(np.abs(b)), 1e-30)
            err = np.max(np.abs(a - b))
            assert err <= TOL["bf16x3"] * scale, (ntau, name, err, scale)


@pytest.mark.parametrize("nao,naux", [(2, 3), (4, 12), (8, 32), (13, 20), (16, 64), (32, 64)])
@pytest.mark.parametrize("prec", ["f64", "f32", "bf16x3"])
def test_sigma_full_solver_parity(orc, nao, naux, prec):
    """The full, non-decomposed solver (ned_sigma_gf2_full, molGF2Solver)
    against the oracle's decomposed one: the reference's own cross-check that
    both give the same self-energy (tests/mol2bSolver-test.cpp:229-237)."""
    print(f"[sigma full] nao={nao} naux={naux} prec={prec}", flush=True)
    rng = np.random.default_rng(300 * nao + naux)
    V, B = sym_V(rng, nao, naux)
    nt, ntau = 6, 7
    G = random_green(rng, nt, ntau, nao, scale=0.5)
    with ned.Context(nao, naux, nt, ntau, 1) as c:
        c.set_coulomb(B)
        c.upload(ned.NED_G, G)
        for tstp in (0, 3, nt):
            S_ref = GreenFunc(nt, ntau, nao)
            orc.sigma_solve(V, tstp, S_ref, G)
            c.sigma_full(tstp, prec)
            S = c.download(ned.NED_SIGMA)
            for name, a, b in (("ret", S.ret_row(tstp), S_ref.ret_row(tstp)),
                               ("les", S.les_col(tstp), S_ref.les_col(tstp)),
                               ("tv", S.tv[tstp], S_ref.tv[tstp])):
                scale = max(np.max(np.abs(b)), 1e-30)
                err = np.max(np.abs(a - b))
                assert err <= TOL[prec] * scale, (name, tstp, err, scale)


def test_sigma_full_solver_chunking_and_reupload(orc, monkeypatch):
    """The full solver in chunks of 3 items equals the whole slice bitwise
    (FP64); after a second ned_set_coulomb its four-index operands are rebuilt
    from the new factor."""
    rng = np.random.default_rng(21)
    nao, naux, nt, ntau = 8, 20, 5, 6
    V, B = sym_V(rng, nao, naux)
    V2, B2 = sym_V(rng, nao, naux)
    G = random_green(rng, nt, ntau, nao, scale=0.5)
    outs = []
    for chunk in (0, 3):
        monkeypatch.setenv("NED_SIGMA_CHUNK", str(chunk))
        with ned.Context(nao, naux, nt, ntau, 1) as c:
            c.set_coulomb(B)
            c.upload(ned.NED_G, G)
            c.sigma_full(nt, "f64")
            outs.append(c.download(ned.NED_SIGMA))
            if chunk == 3:
                c.set_coulomb(B2)
                c.sigma_full(nt, "f64")
                S2 = c.download(ned.NED_SIGMA)
    for comp in ("ret", "les", "tv"):
        assert np.array_equal(getattr(outs[0], comp), getattr(outs[1], comp)), comp
    S_ref = GreenFunc(nt, ntau, nao)
    orc.sigma_solve(V2, nt, S_ref, G)
    assert np.max(np.abs(S2.ret_row(nt) - S_ref.ret_row(nt))) < 1e-12 * np.max(np.abs(S_ref.ret_row(nt)))
