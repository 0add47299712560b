"""Storage modes of the time histories (ned_ctx_create_ex): G^R / G^< kept as
a ring of their last k + 3 rows, and the Sigma^R / Sigma^< triangles in
complex FP32.  The ring changes only where rows live, so every step must
equal the whole-triangle run bit for bit; FP32 Sigma changes only the stored
precision of Sigma, so a Dyson step from the same (FP32-representable) Sigma
must equal the FP64-storage step bit for bit, and a whole run must stay
within the FP32 rounding of Sigma.  (The corrector's update norm sums the
parts of many workgroups with atomics, so it is compared to rounding.)"""
import numpy as np
import pytest

import nedyson_amd as ned
from nedyson_amd import problem

pytestmark = pytest.mark.gpu

RING, S32 = ned.NED_STORE_G_RING, ned.NED_STORE_SIGMA_F32


def _problem(nt=40):
    return problem.Problem(nao=8, naux=24, nelec=4, beta=10.0, ntau=31, nt=nt, dt=0.02, seed=7, spacing=1.0,
                           b_sd=0.1, quench="random", quench_norm=0.3)


def _prepared(p, storage):
    """The free G of the HF Fock matrix on steps 0..k."""
    c = ned.Context(p.nao, p.naux, p.nt, p.ntau, p.k, storage=storage)
    c.set_coulomb(p.B)
    c.write_blocks(ned.NED_G, ned.NED_MAT, 0, p.Gmat)
    c.set_matsubara(p.beta)
    c.g0_from_h0(p.k, p.fock, p.mu, p.beta, p.dt)
    return c


def _booted(p, storage):
    c = _prepared(p, storage)
    c.do_boot(p.fock, p.mu, p.beta, p.dt, "f64", 50, 1e-12)
    return c


def _row(c, T):
    """G^R(T, 0..T) and G^<(0..T, T): time row T of the packed triangles."""
    b = T * (T + 1) // 2
    return (c.read_blocks(ned.NED_G, ned.NED_RET, b, T + 1), c.read_blocks(ned.NED_G, ned.NED_LES, b, T + 1),
            c.read_blocks(ned.NED_G, ned.NED_TV, T * c.np, c.np))


def test_ring_storage_bitwise_equal_to_whole_triangles():
    p = _problem()
    full, ring = _booted(p, 0), _booted(p, RING)
    assert ring.storage() == (RING, 0, p.k)
    for T in range(p.k + 1, p.nt + 1):
        a = full.do_tstp(T, p.h0, p.mu, p.beta, p.dt, "f64", 3, 0.0)
        b = ring.do_tstp(T, p.h0, p.mu, p.beta, p.dt, "f64", 3, 0.0)
        # the update norm sums workgroup parts with atomics: equal to rounding
        assert a[0] == b[0] and abs(a[1] - b[1]) <= 1e-12 * abs(a[1]), (T, a, b)
        for x, y in zip(_row(full, T), _row(ring, T)):
            assert np.array_equal(x, y), T
        assert np.array_equal(full.energy(T, p.h0, p.beta, p.dt), ring.energy(T, p.h0, p.beta, p.dt))
    assert ring.storage() == (RING, p.nt - p.k - 2, p.nt)
    for T in range(p.nt - p.k - 2, p.nt + 1):
        for x, y in zip(_row(full, T), _row(ring, T)):
            assert np.array_equal(x, y), T
    # rows no longer held, and G's whole-history operations, are refused
    with pytest.raises(ned.NedError) as e:
        ring.read_blocks(ned.NED_G, ned.NED_RET, 0, 1)
    assert e.value.code == 1  # NED_ERR_ARG
    with pytest.raises(ned.NedError):
        ring.tti_extend(ned.NED_G, p.nt)
    with pytest.raises(ned.NedError):
        ring.sigma(p.k, "f64")
    full.close()
    ring.close()


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16x3"])
def test_sigma_f32_storage_is_the_rounded_fp64_slice(prec):
    """The slice written into FP32 storage is the one of FP64 storage rounded
    to FP32 (one chunk: every item's retarded value is formed in the
    arithmetic's precision and rounded once)."""
    p = _problem(nt=12)
    a, b = _prepared(p, 0), _prepared(p, S32)
    for c in (a, b):
        c.sigma(p.k, prec)
    n = (p.k + 1) * (p.k + 2) // 2
    for comp in (ned.NED_RET, ned.NED_LES):
        x = a.read_blocks(ned.NED_SIGMA, comp, 0, n)
        y = b.read_blocks(ned.NED_SIGMA, comp, 0, n)
        assert np.array_equal(y, x.astype(np.complex64).astype(np.complex128)), comp
    assert np.array_equal(a.read_blocks(ned.NED_SIGMA, ned.NED_TV, 0, (p.k + 1) * a.np),
                          b.read_blocks(ned.NED_SIGMA, ned.NED_TV, 0, (p.k + 1) * b.np))
    a.close()
    b.close()


@pytest.mark.parametrize("nao", [8, 20])
def test_dyson_step_reads_fp32_sigma_exactly(nao):
    """Every Dyson kernel that reads the Sigma triangles, in its FP32-storage
    instance: a step (and the energy) from Sigma values that FP32 holds
    exactly equals the FP64-storage step bit for bit."""
    p = problem.Problem(nao=nao, naux=3 * nao, nelec=nao // 2, beta=10.0, ntau=31, nt=30, dt=0.02, seed=11,
                        spacing=1.0, b_sd=0.1, quench="random", quench_norm=0.3)
    a, b = _prepared(p, 0), _prepared(p, S32)
    T = 24
    rng = np.random.default_rng(5)
    ntri = (T + 1) * (T + 2) // 2
    for comp in (ned.NED_RET, ned.NED_LES):
        S = (0.05 * (rng.normal(size=(ntri, nao, nao)) + 1j * rng.normal(size=(ntri, nao, nao))))
        S = S.astype(np.complex64).astype(np.complex128)
        a.write_blocks(ned.NED_SIGMA, comp, 0, S)
        b.write_blocks(ned.NED_SIGMA, comp, 0, S)
    for c in (a, b):
        c.g0_from_h0(T, p.fock, p.mu, p.beta, p.dt)
        c.write_hmf(np.repeat(p.fock[None].astype(complex), T + 1, axis=0))
    ea = a.dyson_step(T, p.mu, p.beta, p.dt)
    eb = b.dyson_step(T, p.mu, p.beta, p.dt)
    assert np.allclose(ea, eb, rtol=1e-12, atol=0)  # update norms: atomics, equal to rounding
    for x, y in zip(_row(a, T), _row(b, T)):
        assert np.array_equal(x, y)
    assert np.array_equal(a.energy(T, p.h0, p.beta, p.dt), b.energy(T, p.h0, p.beta, p.dt))
    a.close()
    b.close()


def test_fp32_sigma_storage_run_close_to_fp64():
    """A propagated run with FP64 arithmetic and FP32 Sigma storage stays
    within 1e-6 of the FP64-storage run, and ring + FP32 storage equals FP32
    storage alone bit for bit."""
    p = _problem()
    ref, s32, both = _booted(p, 0), _booted(p, S32), _booted(p, S32 | RING)
    worst = 0.0
    for T in range(p.k + 1, p.nt + 1):
        for c in (ref, s32, both):
            c.do_tstp(T, p.h0, p.mu, p.beta, p.dt, "f64", 3, 0.0)
        r, s, w = _row(ref, T), _row(s32, T), _row(both, T)
        worst = max(worst, max(float(np.max(np.abs(x - y))) for x, y in zip(r, s)))
        for x, y in zip(s, w):
            assert np.array_equal(x, y), T
    print(f"[storage] FP32 Sigma storage: max |dG| over {p.nt - p.k} steps {worst:.2e}", flush=True)
    assert worst < 1e-6
    for c in (ref, s32, both):
        c.close()
