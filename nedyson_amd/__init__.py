"""nedyson_amd: NEdyson's Kadanoff-Baym GF2 timestep on one MI355X (gfx950).

Python binding of the C ABI in include/nedyson.h (libnedyson_amd.so, built
in-tree by ``make -C nedyson_amd/csrc``).  Importing this package does not
touch the GPU; creating a :class:`Context` does.  There is no CPU fallback:
if the library is missing or no HIP device is present, the calls raise.

The classes :class:`molGF2SolverDecomp` and :class:`dyson` keep the call
signatures of the reference (src/mol2bSolver.h:125, src/dyson.h:86-87,114) on
host :class:`~nedyson_amd.green.GreenFunc` objects; :class:`Context` is the
device-resident form used by the timestepping driver and the benchmark.
"""
import ctypes as C
import os

import numpy as np

from .green import GreenFunc  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libnedyson_amd.so")

NED_PREC = {"f64": 0, "f32": 1, "bf16x3": 2}
NED_G, NED_SIGMA = 0, 1
NED_RET, NED_LES, NED_TV, NED_MAT = 0, 1, 2, 3
NED_MAT_HF, NED_MAT_GF2 = 0, 1
_STATUS = {0: "NED_OK", 1: "NED_ERR_ARG", 2: "NED_ERR_NOMEM", 3: "NED_ERR_DEVICE", 4: "NED_ERR_NONFINITE",
           5: "NED_ERR_NOCONV", 6: "NED_ERR_STATE"}

# every symbol declared in include/nedyson.h
EXPORTS = ["ned_last_error", "ned_ctx_create", "ned_ctx_create_ex", "ned_ctx_get_storage", "ned_ctx_destroy", "ned_sync", "ned_ctx_get_dims", "ned_set_coulomb", "ned_green_write",
           "ned_green_read", "ned_hmf_write", "ned_hmf_read", "ned_sigma_gf2_decomp", "ned_sigma_gf2_full",
           "ned_sigma_gf2_decomp_noalloc",
           "ned_sigma_noalloc_scratch_bytes", "ned_hf_decomp",
           "ned_set_matsubara", "ned_matsubara_gf2", "ned_dyson_step", "ned_extrapolate", "ned_dyson_start", "ned_do_boot", "ned_do_tstp",
           "ned_energy", "ned_g0_from_h0", "ned_tti_extend", "ned_profile", "ned_profile_read", "ned_gll_nodes",
           "ned_integrator", "ned_mlp_create", "ned_mlp_destroy", "ned_mlp_nparams", "ned_mlp_get_params",
           "ned_mlp_set_params", "ned_mlp_loss_grad", "ned_mlp_adam_step", "ned_mlp_train", "ned_mlp_predict",
           "ned_mlp_set_bf16", "ned_set_predictor", "ned_predictor_residual", "ned_predictor_stats", "ned_batch_create",
           "ned_batch_destroy", "ned_batch_do_tstp", "ned_gemm_variant_count", "ned_gemm_variant_name",
           "ned_set_graph", "ned_batch_set_grouped"]


class NedError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_STATUS.get(code, code)}: {msg}")
        self.code = code


class _Dims(C.Structure):
    _fields_ = [("nao", C.c_int), ("naux", C.c_int), ("nt", C.c_int), ("ntau", C.c_int), ("k", C.c_int),
                ("sig", C.c_int)]


_lib = None
_vp = C.c_void_p
_dp = C.POINTER(C.c_double)


def build():
    """Compile libnedyson_amd.so for gfx950 (hipcc)."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", os.path.join(_HERE, "csrc"), "-j8"])


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise ImportError(f"{LIBPATH} is not built; run nedyson_amd.build() (hipcc, gfx950)")
    L = C.CDLL(LIBPATH)
    L.ned_last_error.restype = C.c_char_p
    L.ned_ctx_create.argtypes = [C.POINTER(_vp), C.POINTER(_Dims), C.c_int]
    L.ned_ctx_create_ex.argtypes = [C.POINTER(_vp), C.POINTER(_Dims), C.c_int, C.c_uint]
    L.ned_ctx_get_storage.argtypes = [_vp, C.POINTER(C.c_uint), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.ned_ctx_destroy.argtypes = [_vp]
    L.ned_sync.argtypes = [_vp]
    L.ned_set_coulomb.argtypes = [_vp, _dp]
    L.ned_green_write.argtypes = [_vp, C.c_int, C.c_int, C.c_longlong, C.c_longlong, _vp]
    L.ned_green_read.argtypes = [_vp, C.c_int, C.c_int, C.c_longlong, C.c_longlong, _vp]
    L.ned_hmf_write.argtypes = [_vp, C.c_int, C.c_int, _vp]
    L.ned_hmf_read.argtypes = [_vp, C.c_int, C.c_int, _vp]
    L.ned_sigma_gf2_decomp.argtypes = [_vp, C.c_int, C.c_int]
    L.ned_sigma_gf2_decomp_noalloc.argtypes = [_vp, C.c_int, C.c_int]
    L.ned_sigma_gf2_full.argtypes = [_vp, C.c_int, C.c_int]
    L.ned_sigma_noalloc_scratch_bytes.restype = C.c_size_t
    L.ned_hf_decomp.argtypes = [_vp, C.c_int, _dp]
    L.ned_set_matsubara.argtypes = [_vp, C.c_double]
    L.ned_matsubara_gf2.argtypes = [_vp, _dp, C.c_double, C.c_double, C.c_int, C.c_int, _dp, C.c_int, C.c_double,
                                    _dp, C.POINTER(C.c_int), _dp]
    L.ned_dyson_step.argtypes = [_vp, C.c_int, C.c_double, C.c_double, C.c_double, _dp]
    L.ned_extrapolate.argtypes = [_vp, C.c_int]
    L.ned_dyson_start.argtypes = [_vp, C.c_double, C.c_double, C.c_double, _dp]
    L.ned_do_boot.argtypes = [_vp, _dp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double,
                              C.POINTER(C.c_int), _dp]
    L.ned_set_graph.argtypes = [_vp, C.c_int]
    L.ned_do_tstp.argtypes = [_vp, C.c_int, _dp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double,
                              C.POINTER(C.c_int), _dp]
    L.ned_energy.argtypes = [_vp, C.c_int, _dp, C.c_double, C.c_double, _dp]
    L.ned_g0_from_h0.argtypes = [_vp, C.c_int, _dp, C.c_double, C.c_double, C.c_double]
    L.ned_tti_extend.argtypes = [_vp, C.c_int, C.c_int]
    L.ned_profile.argtypes = [_vp, C.c_int]
    L.ned_profile_read.argtypes = [_vp, _dp]
    L.ned_gll_nodes.argtypes = [C.c_int, _dp]
    L.ned_integrator.argtypes = [C.c_int] + [_dp] * 7
    _fp = C.POINTER(C.c_float)
    L.ned_mlp_create.argtypes = [C.POINTER(_vp), C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong]
    L.ned_mlp_destroy.argtypes = [_vp]
    L.ned_mlp_nparams.argtypes = [_vp, C.POINTER(C.c_longlong)]
    L.ned_mlp_get_params.argtypes = [_vp, _fp]
    L.ned_mlp_set_params.argtypes = [_vp, _fp]
    L.ned_mlp_loss_grad.argtypes = [_vp, _fp, _fp, C.c_int, _dp, _fp]
    L.ned_mlp_adam_step.argtypes = [_vp, C.c_float]
    L.ned_mlp_train.argtypes = [_vp, _fp, _fp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_ulonglong, _fp]
    L.ned_mlp_predict.argtypes = [_vp, _fp, C.c_int, _fp]
    L.ned_mlp_set_bf16.argtypes = [_vp, C.c_int]
    L.ned_set_predictor.argtypes = [_vp, _vp]
    L.ned_predictor_residual.argtypes = [_vp, C.c_int, _fp]
    L.ned_predictor_stats.argtypes = [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.ned_batch_create.argtypes = [C.POINTER(_vp), C.POINTER(_vp), C.c_int]
    L.ned_batch_set_grouped.argtypes = [_vp, C.c_int]
    L.ned_batch_destroy.argtypes = [_vp]
    L.ned_batch_do_tstp.argtypes = [_vp, C.c_int, C.POINTER(_dp), _dp, _dp, _dp, C.c_int, C.c_int, C.c_double,
                                    C.POINTER(C.c_int), _dp, C.POINTER(C.c_int)]
    for name in EXPORTS[1:]:
        getattr(L, name).restype = C.c_int
    L.ned_gemm_variant_name.argtypes = [C.c_int]
    L.ned_gemm_variant_name.restype = C.c_char_p
    _lib = L
    return L


def _check(code):
    if code != 0:
        raise NedError(code, lib().ned_last_error().decode())


def _cptr(a):
    return a.ctypes.data_as(_vp)


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _c128(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a


def prec_code(prec):
    if isinstance(prec, str):
        return NED_PREC[prec]
    return int(prec)


def gemm_variants():
    """Names of the split-bf16 GEMM tile variants; NED_NT_VARIANT=<index>
    pins one (read on every GEMM call)."""
    L = lib()
    return [L.ned_gemm_variant_name(v).decode() for v in range(L.ned_gemm_variant_count())]


def gll_nodes(ntau):
    x = np.zeros(ntau + 1)
    _check(lib().ned_gll_nodes(ntau, _dptr(x)))
    return x


def integrator(k):
    k1 = k + 1
    arrs = [np.zeros(k1 * k1), np.zeros(k1 * k1), np.zeros(k1 ** 3), np.zeros(k + 2), np.zeros(k1 * k1),
            np.zeros(k1 * k1), np.zeros(k1)]
    _check(lib().ned_integrator(k, *[_dptr(a) for a in arrs]))
    return dict(zip(["poly_interp", "poly_diff", "poly_integ", "bd", "start", "Omega", "ex"], arrs))


NED_STORE_SIGMA_F32 = 1
NED_STORE_G_RING = 2


class Context:
    """One system resident on one GPU (see include/nedyson.h).  storage:
    NED_STORE_SIGMA_F32 and / or NED_STORE_G_RING (ned_ctx_create_ex), 0 for
    whole FP64 histories."""

    def __init__(self, nao, naux, nt, ntau, k, sig=-1, device=0, storage=0):
        self.nao, self.naux, self.nt, self.ntau, self.k, self.sig = nao, naux, nt, ntau, k, sig
        self.np = ntau + 1
        self.h = _vp()
        d = _Dims(nao, naux, nt, ntau, k, sig)
        _check(lib().ned_ctx_create_ex(C.byref(self.h), C.byref(d), device, storage))

    def storage(self):
        """(flags, first G row held, last G row held)"""
        f, a, b = C.c_uint(0), C.c_int(0), C.c_int(0)
        _check(lib().ned_ctx_get_storage(self.h, C.byref(f), C.byref(a), C.byref(b)))
        return f.value, a.value, b.value

    def close(self):
        if getattr(self, "h", None):
            lib().ned_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- data movement
    def set_coulomb(self, B):
        B = np.ascontiguousarray(B, dtype=np.float64)
        assert B.shape == (self.naux, self.nao, self.nao)
        _check(lib().ned_set_coulomb(self.h, _dptr(B)))

    def write_blocks(self, which, comp, first, blocks):
        b = _c128(blocks)
        nb = b.size // (self.nao * self.nao)
        _check(lib().ned_green_write(self.h, which, comp, first, nb, _cptr(b)))

    def read_blocks(self, which, comp, first, nb):
        out = np.zeros((nb, self.nao, self.nao), dtype=np.complex128)
        _check(lib().ned_green_read(self.h, which, comp, first, nb, _cptr(out)))
        return out

    def upload(self, which, g):
        self.write_blocks(which, NED_RET, 0, g.ret)
        self.write_blocks(which, NED_LES, 0, g.les)
        self.write_blocks(which, NED_TV, 0, g.tv)
        self.write_blocks(which, NED_MAT, 0, g.mat)

    def download(self, which, g=None):
        if g is None:
            g = GreenFunc(self.nt, self.ntau, self.nao, self.sig)
        ntri = (self.nt + 1) * (self.nt + 2) // 2
        g.ret[:] = self.read_blocks(which, NED_RET, 0, ntri)
        g.les[:] = self.read_blocks(which, NED_LES, 0, ntri)
        g.tv[:] = self.read_blocks(which, NED_TV, 0, (self.nt + 1) * self.np).reshape(g.tv.shape)
        g.mat[:] = self.read_blocks(which, NED_MAT, 0, self.np)
        return g

    def write_hmf(self, hmf, first=0):
        h = _c128(hmf)
        _check(lib().ned_hmf_write(self.h, first, h.shape[0], _cptr(h)))

    def read_hmf(self, first=0, count=None):
        count = self.nt + 1 - first if count is None else count
        out = np.zeros((count, self.nao, self.nao), dtype=np.complex128)
        _check(lib().ned_hmf_read(self.h, first, count, _cptr(out)))
        return out

    # ---- hot path
    def sigma(self, tstp, prec="f64"):
        _check(lib().ned_sigma_gf2_decomp(self.h, tstp, prec_code(prec)))

    def sigma_full(self, tstp, prec="f64"):
        """molGF2Solver::solve: the same slice of Sigma from the full
        (non-decomposed) four-index interaction."""
        _check(lib().ned_sigma_gf2_full(self.h, tstp, prec_code(prec)))

    def sigma_noalloc(self, tstp, prec="f64"):
        """ned_sigma_gf2_decomp_noalloc: the same slice of Sigma with no device
        allocation (static per-GPU scratch)."""
        _check(lib().ned_sigma_gf2_decomp_noalloc(self.h, tstp, prec_code(prec)))

    def hf(self, tstp, h0):
        h0 = np.ascontiguousarray(h0, dtype=np.float64)
        _check(lib().ned_hf_decomp(self.h, tstp, _dptr(h0)))

    def set_matsubara(self, beta):
        _check(lib().ned_set_matsubara(self.h, beta))

    def matsubara_gf2(self, h1, beta, nelec=None, mu=None, mode="gf2", maxiter=200, tol=1e-10):
        """Equilibrium G^M (GF2 or HF) on the GLL nodes, written into the
        context (see ned_matsubara_gf2).  With nelec, mu is solved for (mu is
        then only the starting value); without it, mu is held fixed.
        Returns (mu, fock, iterations, last update)."""
        h1 = np.ascontiguousarray(h1, dtype=np.float64)
        fix = nelec is None
        if fix and mu is None:
            raise ValueError("give nelec or a fixed mu")
        m = C.c_double(0.0 if mu is None else float(mu))
        fock = np.zeros((self.nao, self.nao))
        it, err = C.c_int(0), C.c_double(0.0)
        _check(lib().ned_matsubara_gf2(self.h, _dptr(h1), 0.0 if fix else float(nelec), float(beta),
                                       NED_MAT_GF2 if mode == "gf2" else NED_MAT_HF, int(fix), C.byref(m), maxiter,
                                       tol, _dptr(fock), C.byref(it), C.byref(err)))
        return m.value, fock, it.value, err.value

    def dyson_step(self, tstp, mu, beta, dt):
        e = np.zeros(3)
        _check(lib().ned_dyson_step(self.h, tstp, mu, beta, dt, _dptr(e)))
        return e

    def extrapolate(self, tstp):
        _check(lib().ned_extrapolate(self.h, tstp))

    def dyson_start(self, mu, beta, dt):
        e = np.zeros(1)
        _check(lib().ned_dyson_start(self.h, mu, beta, dt, _dptr(e)))
        return float(e[0])

    def do_boot(self, fock, mu, beta, dt, prec="f64", maxiter=100, tol=1e-10):
        fock = np.ascontiguousarray(fock, dtype=np.float64)
        it = C.c_int(0)
        e = np.zeros(1)
        _check(lib().ned_do_boot(self.h, _dptr(fock), mu, beta, dt, prec_code(prec), maxiter, tol, C.byref(it),
                                 _dptr(e)))
        return it.value, float(e[0])

    def set_graph(self, enable=True):
        """ned_set_graph: replay the corrector passes after the first of every
        step from one captured HIP graph."""
        _check(lib().ned_set_graph(self.h, int(bool(enable))))

    def do_tstp(self, tstp, h0, mu, beta, dt, prec="f64", corr_max=3, corr_tol=0.0):
        h0 = np.ascontiguousarray(h0, dtype=np.float64)
        it = C.c_int(0)
        e = np.zeros(1)
        _check(lib().ned_do_tstp(self.h, tstp, _dptr(h0), mu, beta, dt, prec_code(prec), corr_max, corr_tol,
                                 C.byref(it), _dptr(e)))
        return it.value, float(e[0])

    def energy(self, tstp, h0, beta, dt):
        h0 = np.ascontiguousarray(h0, dtype=np.float64)
        out = np.zeros(3)
        _check(lib().ned_energy(self.h, tstp, _dptr(h0), beta, dt, _dptr(out)))
        return out

    def g0_from_h0(self, tmax, fock, mu, beta, dt):
        fock = np.ascontiguousarray(fock, dtype=np.float64)
        _check(lib().ned_g0_from_h0(self.h, tmax, _dptr(fock), mu, beta, dt))

    def tti_extend(self, which, tsrc):
        _check(lib().ned_tti_extend(self.h, which, tsrc))

    def set_predictor(self, mlp):
        """Use `mlp` (an Mlp with 768 inputs, 128 outputs) for the initial guess
        of do_tstp, or the polynomial predictor alone when None."""
        _check(lib().ned_set_predictor(self.h, mlp.h if mlp is not None else None))
        self._pred = mlp  # keep the network alive while the context uses it

    def predictor_residual(self, tstp):
        out = np.zeros((self.nao * self.nao, PRED_OUTPUTS), dtype=np.float32)
        _check(lib().ned_predictor_residual(self.h, tstp, _fptr(out)))
        return out

    def predictor_stats(self):
        u, f = C.c_longlong(0), C.c_longlong(0)
        _check(lib().ned_predictor_stats(self.h, C.byref(u), C.byref(f)))
        return u.value, f.value

    def profile(self, enable=True):
        _check(lib().ned_profile(self.h, int(enable)))

    def profile_read(self):
        out = np.zeros(8)
        _check(lib().ned_profile_read(self.h, _dptr(out)))
        return {"ugemm_ms": out[0], "ugemm_n": int(out[1]), "sigma_ms": out[2], "sigma_n": int(out[3]),
                "dyson_ms": out[4], "dyson_n": int(out[5]), "ugemm_flops": out[6]}

    def sync(self):
        _check(lib().ned_sync(self.h))


# predictor network shape: residuals of the last six steps at 2 x 32 points
# (real, imaginary) per orbital pair in, the next step's out
PRED_OUTPUTS = 128
PRED_INPUTS = 6 * PRED_OUTPUTS


def _fptr(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Mlp:
    """The predictor network on one GPU (ned_mlp_*, include/nedyson.h)."""

    def __init__(self, n_in=PRED_INPUTS, width=512, n_out=PRED_OUTPUTS, device=0, seed=0):
        self.n_in, self.width, self.n_out = n_in, width, n_out
        self.h = _vp()
        _check(lib().ned_mlp_create(C.byref(self.h), n_in, width, n_out, device, seed))
        n = C.c_longlong(0)
        _check(lib().ned_mlp_nparams(self.h, C.byref(n)))
        self.nparams = n.value

    def close(self):
        if getattr(self, "h", None):
            lib().ned_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        out = np.zeros(self.nparams, dtype=np.float32)
        _check(lib().ned_mlp_get_params(self.h, _fptr(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        _check(lib().ned_mlp_set_params(self.h, _fptr(p)))

    def split(self, flat):
        """Flat parameters -> [W1, b1, W2, b2, W3, b3]."""
        shapes = [(self.width, self.n_in), (self.width,), (self.width, self.width), (self.width,),
                  (self.n_out, self.width), (self.n_out,)]
        out, o = [], 0
        for sh in shapes:
            n = int(np.prod(sh))
            out.append(flat[o:o + n].reshape(sh))
            o += n
        return out

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = C.c_double(0)
        _check(lib().ned_mlp_loss_grad(self.h, _fptr(X), _fptr(T), X.shape[0], C.byref(loss), _fptr(g)))
        return loss.value, g

    def adam_step(self, lr):
        _check(lib().ned_mlp_adam_step(self.h, lr))

    def train(self, X, T, epochs=10, batch=1024, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        losses = np.zeros(epochs, dtype=np.float32)
        _check(lib().ned_mlp_train(self.h, _fptr(X), _fptr(T), X.shape[0], epochs, batch, lr, seed, _fptr(losses)))
        return losses

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], self.n_out), dtype=np.float32)
        _check(lib().ned_mlp_predict(self.h, _fptr(X), X.shape[0], _fptr(Y)))
        return Y

    def set_bf16(self, on=True):
        """Inference (predict and the timestep's guess) with bf16 weights and
        activations on the bf16 matrix cores, FP32 accumulation; training stays FP32."""
        _check(lib().ned_mlp_set_bf16(self.h, 1 if on else 0))


class Batch:
    """Independent systems stepped together (ned_batch_*): grouped (the Dyson
    recurrences of all systems as one launch per link) by default, or one host
    thread and stream per system with grouped=False."""

    def __init__(self, contexts, grouped=True):
        self.contexts = list(contexts)
        arr = (_vp * len(self.contexts))(*[c.h for c in self.contexts])
        self.h = _vp()
        _check(lib().ned_batch_create(C.byref(self.h), arr, len(self.contexts)))
        _check(lib().ned_batch_set_grouped(self.h, int(bool(grouped))))

    def close(self):
        if getattr(self, "h", None):
            lib().ned_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def do_tstp(self, tstp, h0s, mus, betas, dts, prec="f64", corr_max=3, corr_tol=0.0):
        n = len(self.contexts)
        hs = [np.ascontiguousarray(h, dtype=np.float64) for h in h0s]
        hp = (_dp * n)(*[_dptr(h) for h in hs])
        mu = np.ascontiguousarray(mus, dtype=np.float64)
        be = np.ascontiguousarray(betas, dtype=np.float64)
        dt = np.ascontiguousarray(dts, dtype=np.float64)
        it = (C.c_int * n)()
        err = np.zeros(n)
        st = (C.c_int * n)()
        _check(lib().ned_batch_do_tstp(self.h, tstp, hp, _dptr(mu), _dptr(be), _dptr(dt), prec_code(prec), corr_max,
                                       corr_tol, it, _dptr(err), st))
        return list(it), err


class molGF2SolverDecomp:
    """Reference-signature wrapper: molGF2SolverDecomp(Vija, Viaj).solve(tstp, Sigma, G)
    (src/mol2bSolver.h:101-125) on host GreenFunc objects, computed on the GPU."""

    def __init__(self, Vija, Viaj=None, prec="f64", device=0):
        V = np.ascontiguousarray(Vija, dtype=np.float64)
        self.nao, self.naux = V.shape[0], V.shape[2]
        self.B = np.ascontiguousarray(np.transpose(V, (2, 0, 1)))
        self.prec, self.device = prec, device
        self._ctx = None

    def _context(self, G):
        c = self._ctx
        if c is None or (c.nt, c.ntau) != (G.nt, G.ntau):
            if c is not None:
                c.close()
            c = Context(self.nao, self.naux, G.nt, G.ntau, min(5, max(1, G.nt)), G.sig, self.device)
            c.set_coulomb(self.B)
            self._ctx = c
        return c

    def solve(self, tstp, Sigma, G):
        c = self._context(G)
        c.upload(NED_G, G)
        c.sigma(tstp, self.prec)
        ntri = (G.nt + 1) * (G.nt + 2) // 2
        s = tstp * (tstp + 1) // 2
        Sigma.ret[s:s + tstp + 1] = c.read_blocks(NED_SIGMA, NED_RET, s, tstp + 1)
        Sigma.les[s:s + tstp + 1] = c.read_blocks(NED_SIGMA, NED_LES, s, tstp + 1)
        Sigma.tv[tstp] = c.read_blocks(NED_SIGMA, NED_TV, tstp * c.np, c.np)
        del ntri


class dyson:
    """Reference-signature wrapper: dyson(nt, ntau, nao, k).dyson_step(n, G, Sig, hmf, mu,
    beta, dt) and Extrapolate(n, G) (src/dyson.h:41-114) on host GreenFunc objects."""

    def __init__(self, nt, ntau, nao, k, device=0):
        self.nt, self.ntau, self.nao, self.k = nt, ntau, nao, k
        self.ctx = Context(nao, 1, nt, ntau, k, -1, device)
        self._beta = None

    def _prep(self, G, Sig, hmf, beta):
        c = self.ctx
        c.upload(NED_G, G)
        if Sig is not None:
            c.upload(NED_SIGMA, Sig)
        if hmf is not None:
            c.write_hmf(hmf)
        if beta is not None:
            c.set_matsubara(beta)

    def dyson_step(self, n, G, Sig, hmf, mu, beta, dt):
        self._prep(G, Sig, hmf, beta)
        e = self.ctx.dyson_step(n, mu, beta, dt)
        self.ctx.download(NED_G, G)
        return e

    def Extrapolate(self, n, G):
        self._prep(G, None, None, None)
        self.ctx.extrapolate(n)
        self.ctx.download(NED_G, G)

    def dyson_start(self, G, Sig, hmf, mu, beta, dt):
        self._prep(G, Sig, hmf, beta)
        e = self.ctx.dyson_start(mu, beta, dt)
        self.ctx.download(NED_G, G)
        return e
