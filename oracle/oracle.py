"""ctypes wrapper of the CPU oracle (liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker.  The product package never
imports it.

Contour functions are passed as any object with integer attributes nt, ntau,
nao, sig and complex128 C-contiguous numpy arrays ret, les (packed triangles,
shape (ntri, nao, nao)), tv (nt+1, ntau+1, nao, nao) and mat (ntau+1, nao, nao)
-- the packing of green_func (src/greens.cpp:177-195).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liborc.so")
_lib = None

_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


class _Green(C.Structure):
    _fields_ = [("nt", C.c_int), ("ntau", C.c_int), ("nao", C.c_int), ("sig", C.c_int),
                ("ret", _vp), ("les", _vp), ("tv", _vp), ("mat", _vp)]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        L = _lib
        L.orc_gregory_table.argtypes = [C.c_int, C.c_int, _dp]
        L.orc_gll.argtypes = [C.c_int, _dp, _dp]
        L.orc_mixing_kernel.argtypes = [C.c_int, C.c_int, _vp, C.c_double, C.c_int, _vp]
        L.orc_integrator.argtypes = [C.c_int] + [_dp] * 7
        for f in ("orc_sigma_solve",):
            getattr(L, f).argtypes = [C.c_int, C.c_int, _dp, C.c_int, C.POINTER(_Green), C.POINTER(_Green), C.c_int]
        L.orc_sigma_solve_loop.argtypes = [C.c_int, C.c_int, _dp, C.c_int, C.POINTER(_Green), C.POINTER(_Green)]
        L.orc_sigma_point.argtypes = [C.c_int, C.c_int, _dp, C.c_int, C.c_int, C.c_int, C.POINTER(_Green),
                                      C.POINTER(_Green)]
        L.orc_hf.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _vp, _vp, C.c_int]
        L.orc_dyson_new.restype = _vp
        L.orc_dyson_new.argtypes = [C.c_int] * 4
        L.orc_dyson_free.argtypes = [_vp]
        PG = C.POINTER(_Green)
        L.orc_dyson_step.restype = C.c_double
        L.orc_dyson_step.argtypes = [_vp, C.c_int, PG, PG, _vp, C.c_double, C.c_double, C.c_double, _dp]
        L.orc_dyson_start.restype = C.c_double
        L.orc_dyson_start.argtypes = [_vp, PG, PG, _vp, C.c_double, C.c_double, C.c_double]
        L.orc_extrapolate.argtypes = [_vp, C.c_int, PG]
        L.orc_energy_conv.restype = C.c_double
        L.orc_energy_conv.argtypes = [_vp, C.c_int, PG, PG, C.c_double, C.c_double]
        L.orc_ctv1.argtypes = [_vp, _vp, PG, PG, PG, C.c_int, C.c_double]
        L.orc_ctv_tstp.argtypes = [_vp, C.c_int, PG, PG, PG, PG, PG, C.c_double, C.c_double]
        L.orc_cles2.argtypes = [_vp, PG, PG, PG, PG, C.c_int, C.c_double, _vp]
        L.orc_cles3.argtypes = [_vp, PG, PG, PG, PG, C.c_int, C.c_double, _vp]
        L.orc_conv_mixing.argtypes = [_vp, _vp, _vp, PG, C.c_double]
        L.orc_g0_from_h0.argtypes = [_vp, PG, C.c_double, _dp, C.c_double, C.c_double]
        L.orc_g0_matsubara.argtypes = [_vp, PG, C.c_double, _dp, C.c_double]
        L.orc_do_boot.restype = C.c_int
        L.orc_do_boot.argtypes = [_vp, C.c_int, _dp, PG, PG, _vp, _dp, C.c_double, C.c_double, C.c_double, C.c_int,
                                  C.c_double, _dp, C.c_int]
        L.orc_do_tstp.restype = C.c_int
        L.orc_do_tstp.argtypes = [_vp, C.c_int, _dp, C.c_int, PG, PG, _vp, _dp, C.c_double, C.c_double, C.c_double,
                                  C.c_int, C.c_double, _dp, C.c_int, C.c_int]
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_vp) if a is not None else None


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _check(a, dtype):
    assert a.dtype == dtype and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)


def gview(g):
    for name in ("ret", "les", "tv", "mat"):
        _check(getattr(g, name), np.complex128)
    s = _Green(g.nt, g.ntau, g.nao, g.sig, _ptr(g.ret), _ptr(g.les), _ptr(g.tv), _ptr(g.mat))
    s._keep = g
    return C.byref(s)


def integrator(k):
    k1 = k + 1
    arrs = [np.zeros(k1 * k1), np.zeros(k1 * k1), np.zeros(k1 ** 3), np.zeros(k + 2), np.zeros(k1 * k1),
            np.zeros(k1 * k1), np.zeros(k1)]
    lib().orc_integrator(k, *[_dptr(a) for a in arrs])
    names = ["poly_interp", "poly_diff", "poly_integ", "bd", "start", "Omega", "ex"]
    return dict(zip(names, arrs))


def gregory_table(k, nmax):
    w = np.zeros((nmax + 1) * (nmax + 1))
    lib().orc_gregory_table(k, nmax, _dptr(w))
    return w.reshape(nmax + 1, nmax + 1)


def gll(ntau):
    x = np.zeros(ntau + 1)
    m = np.zeros((ntau + 1) * (ntau + 1))
    lib().orc_gll(ntau, _dptr(x), _dptr(m))
    return x, m.reshape(ntau + 1, ntau + 1)


def mixing_kernel(ntau, B, beta, sig):
    B = np.ascontiguousarray(B, dtype=np.complex128)
    nao = B.shape[1]
    K = np.zeros((ntau + 1, ntau + 1, nao, nao), dtype=np.complex128)
    lib().orc_mixing_kernel(ntau, nao, _ptr(B), beta, sig, _ptr(K))
    return K


def sigma_solve(V, tstp, S, G, nthreads=1):
    V = np.ascontiguousarray(V, dtype=np.float64)
    lib().orc_sigma_solve(V.shape[0], V.shape[2], _dptr(V), tstp, gview(S), gview(G), nthreads)


def sigma_solve_loop(V, tstp, S, G):
    V = np.ascontiguousarray(V, dtype=np.float64)
    lib().orc_sigma_solve_loop(V.shape[0], V.shape[2], _dptr(V), tstp, gview(S), gview(G))


def sigma_point(V, comp, tstp, t, S, G):
    V = np.ascontiguousarray(V, dtype=np.float64)
    lib().orc_sigma_point(V.shape[0], V.shape[2], _dptr(V), comp, tstp, t, gview(S), gview(G))


def hf(V, tstp, hmf, rho, loop=False):
    V = np.ascontiguousarray(V, dtype=np.float64)
    _check(hmf, np.complex128)
    rho = np.ascontiguousarray(rho, dtype=np.complex128)
    lib().orc_hf(V.shape[0], V.shape[2], _dptr(V), tstp, _ptr(hmf), _ptr(rho), int(loop))


class Dyson:
    """The reference's `dyson` object (src/dyson.h:20-140) on the CPU."""

    def __init__(self, nt, ntau, nao, k):
        self.nt, self.ntau, self.nao, self.k = nt, ntau, nao, k
        self.h = lib().orc_dyson_new(nt, ntau, nao, k)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_dyson_free(self.h)
            self.h = None

    def dyson_step(self, n, G, S, hmf, mu, beta, dt):
        errs = np.zeros(3)
        _check(hmf, np.complex128)
        e = lib().orc_dyson_step(self.h, n, gview(G), gview(S), _ptr(hmf), mu, beta, dt, _dptr(errs))
        return e, errs

    def dyson_start(self, G, S, hmf, mu, beta, dt):
        return lib().orc_dyson_start(self.h, gview(G), gview(S), _ptr(hmf), mu, beta, dt)

    def extrapolate(self, n, G):
        lib().orc_extrapolate(self.h, n, gview(G))

    def energy_conv(self, tstp, S, G, beta, dt):
        return lib().orc_energy_conv(self.h, tstp, gview(S), gview(G), beta, dt)

    def ctv1(self, A, Acc, B, n, dt):
        res = np.zeros((self.ntau + 1, self.nao, self.nao), dtype=np.complex128)
        lib().orc_ctv1(self.h, _ptr(res), gview(A), gview(Acc), gview(B), n, dt)
        return res

    def ctv_tstp(self, tstp, C_, A, Acc, B, Bcc, beta, dt):
        lib().orc_ctv_tstp(self.h, tstp, gview(C_), gview(A), gview(Acc), gview(B), gview(Bcc), beta, dt)

    def cles2(self, A, Acc, B, Bcc, m, dt):
        res = np.zeros((self.nt + 1, self.nao, self.nao), dtype=np.complex128)
        lib().orc_cles2(self.h, gview(A), gview(Acc), gview(B), gview(Bcc), m, dt, _ptr(res))
        return res

    def cles3(self, A, Acc, B, Bcc, m, beta):
        res = np.zeros((self.nt + 1, self.nao, self.nao), dtype=np.complex128)
        lib().orc_cles3(self.h, gview(A), gview(Acc), gview(B), gview(Bcc), m, beta, _ptr(res))
        return res

    def conv_mixing(self, A, B, beta):
        A = np.ascontiguousarray(A, dtype=np.complex128)
        C_ = np.zeros_like(A)
        lib().orc_conv_mixing(self.h, _ptr(C_), _ptr(A), gview(B), beta)
        return C_

    def g0_from_h0(self, G, mu, H0, beta, h):
        H0 = np.ascontiguousarray(H0, dtype=np.float64)
        lib().orc_g0_from_h0(self.h, gview(G), mu, _dptr(H0), beta, h)

    def g0_matsubara(self, G, mu, H0, beta):
        H0 = np.ascontiguousarray(H0, dtype=np.float64)
        lib().orc_g0_matsubara(self.h, gview(G), mu, _dptr(H0), beta)

    def do_boot(self, V, G, S, hmf, fock, mu, beta, dt, maxiter, tol, nthreads=1):
        V = np.ascontiguousarray(V, dtype=np.float64)
        fock = np.ascontiguousarray(fock, dtype=np.float64)
        err = np.zeros(1)
        it = lib().orc_do_boot(self.h, V.shape[2], _dptr(V), gview(G), gview(S), _ptr(hmf), _dptr(fock), mu, beta,
                               dt, maxiter, tol, _dptr(err), nthreads)
        return it, float(err[0])

    def do_tstp(self, V, tstp, G, S, hmf, h0, mu, beta, dt, corr_max, corr_tol=0.0, nthreads=1, extrapolate=True):
        V = np.ascontiguousarray(V, dtype=np.float64)
        h0 = np.ascontiguousarray(h0, dtype=np.float64)
        err = np.zeros(1)
        it = lib().orc_do_tstp(self.h, V.shape[2], _dptr(V), tstp, gview(G), gview(S), _ptr(hmf), _dptr(h0), mu,
                               beta, dt, corr_max, corr_tol, _dptr(err), nthreads, int(extrapolate))
        return it, float(err[0])
