"""ned_sigma_gf2_decomp_noalloc: the self-energy slice with no device
allocation (chunk scratch in a static per-GPU array) against
ned_sigma_gf2_decomp on the same context, through the C ABI."""
import numpy as np
import pytest

import nedyson_amd as ned
from helpers import random_green, sym_V

pytestmark = pytest.mark.gpu


def _both(B, G, tstp, prec):
    """Sigma slice tstp by the allocating entry point, then by the
    no-allocation one on the same context (the slice is overwritten)."""
    naux, nao = B.shape[0], B.shape[1]
    with ned.Context(nao, naux, G.nt, G.ntau, 1) as c:
        c.set_coulomb(B)
        c.upload(ned.NED_G, G)
        c.sigma(tstp, prec)
        ref = c.download(ned.NED_SIGMA)
        c.upload(ned.NED_SIGMA, ned.GreenFunc(G.nt, G.ntau, nao))
        c.sigma_noalloc(tstp, prec)
        out = c.download(ned.NED_SIGMA)
    return ref, out


def _same(ref, out, tstp):
    for name, a, b in (("ret", out.ret_row(tstp), ref.ret_row(tstp)),
                       ("les", out.les_col(tstp), ref.les_col(tstp)),
                       ("tv", out.tv[tstp], ref.tv[tstp])):
        assert np.array_equal(a, b), (name, np.max(np.abs(a - b)))


def test_noalloc_matches_allocating_entry_point(monkeypatch):
    """Two calls in a row: nao = 64, naux = 256 (16 items, chunked 9 + 7 in the
    256 MiB scratch in bf16x3 with one tile variant pinned, 4 x 4 in FP64;
    bitwise), then a smaller nao = 16, naux = 64 slice in FP64 and bf16x3
    (bitwise)."""
    assert lib_scratch() == 256 << 20
    rng = np.random.default_rng(11)
    monkeypatch.setenv("NED_NT_VARIANT", "10")  # 256 x 256 fits every M = 2 naux P here
    V, B = sym_V(rng, 64, 256)
    G = random_green(rng, 3, 7, 64, scale=0.2)
    for prec in ("bf16x3", "f64"):  # FP64: chunks of 4 against one of 16
        print(f"[noalloc] nao=64 naux=256 tstp=3 {prec}", flush=True)
        ref, out = _both(B, G, 3, prec)
        _same(ref, out, 3)
    V, B = sym_V(rng, 16, 64)
    G = random_green(rng, 5, 6, 16, scale=0.5)
    for prec in ("f64", "bf16x3"):
        print(f"[noalloc] nao=16 naux=64 tstp=2 {prec}", flush=True)
        ref, out = _both(B, G, 2, prec)
        _same(ref, out, 2)


def test_noalloc_rejects_oversized_shape():
    """nao = 64, naux = 2048 in FP64: one time point needs ~400 MB of chunk
    scratch, more than the static array holds: NED_ERR_ARG, nothing run."""
    rng = np.random.default_rng(12)
    V, B = sym_V(rng, 64, 2048)
    G = random_green(rng, 1, 5, 64, scale=0.2)
    print("[noalloc] oversized nao=64 naux=2048", flush=True)
    with ned.Context(64, 2048, 1, 5, 1) as c:
        c.set_coulomb(B)
        c.upload(ned.NED_G, G)
        with pytest.raises(ned.NedError) as e:
            c.sigma_noalloc(1, "f64")
        assert e.value.code == 1


def lib_scratch():
    return int(ned.lib().ned_sigma_noalloc_scratch_bytes())
