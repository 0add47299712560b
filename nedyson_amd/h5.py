"""GREEN files in the reference's HDF5 layout (green_func::print_to_file /
read_from_file, src/greens.cpp:199-306), through libnedyson_h5.so: groups
with the flat complex datasets GR, GL, GTV, GM and the int scalars nt, ntau,
sig, nao.  Host GreenFunc objects and device contexts."""
import ctypes as C
import os

import numpy as np

from . import NED_G, NED_SIGMA, lib as _main_lib
from .green import GreenFunc

H5PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libnedyson_h5.so")
_h5 = None


class H5Error(RuntimeError):
    pass


def lib():
    global _h5
    if _h5 is not None:
        return _h5
    if not os.path.exists(H5PATH):
        raise ImportError(f"{H5PATH} is not built (needs HDF5 headers; see nedyson_amd/csrc/Makefile)")
    _main_lib()
    L = C.CDLL(H5PATH)
    vp, ip = C.c_void_p, C.POINTER(C.c_int)
    L.ned_h5_last_error.restype = C.c_char_p
    L.ned_h5_green_dims.argtypes = [C.c_char_p, C.c_char_p, ip, ip, ip, ip]
    L.ned_h5_green_read_host.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp]
    L.ned_h5_green_write_host.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp]
    L.ned_h5_green_read.argtypes = [vp, C.c_int, C.c_char_p, C.c_char_p]
    L.ned_h5_green_write.argtypes = [vp, C.c_int, C.c_char_p, C.c_char_p, C.c_int]
    _h5 = L
    return L


def _check(code):
    if code != 0:
        raise H5Error(lib().ned_h5_last_error().decode())


def _b(s):
    return str(s).encode()


def dims(path, group=""):
    """(nt, ntau, nao, sig) of the GREEN stored in group."""
    v = [C.c_int(0) for _ in range(4)]
    _check(lib().ned_h5_green_dims(_b(path), _b(group), *[C.byref(x) for x in v]))
    return tuple(x.value for x in v)


def read_green(path, group=""):
    """green_func::read_from_file into a host GreenFunc."""
    nt, ntau, nao, sig = dims(path, group)
    g = GreenFunc(nt, ntau, nao, sig)
    _check(lib().ned_h5_green_read_host(_b(path), _b(group), nt, ntau, nao, g.ret.ctypes.data, g.les.ctypes.data,
                                        g.tv.ctypes.data, g.mat.ctypes.data))
    return g


def write_green(g, path, group=""):
    """green_func::print_to_file of a host GreenFunc (GM for the Matsubara part)."""
    arrs = [np.ascontiguousarray(a, dtype=np.complex128) for a in (g.ret, g.les, g.tv, g.mat)]
    _check(lib().ned_h5_green_write_host(_b(path), _b(group), g.nt, g.ntau, g.nao, g.sig,
                                         *[a.ctypes.data for a in arrs]))


def read_context(ctx, path, group="", which=NED_G):
    """read_from_file into G (NED_G) or Sigma (NED_SIGMA) of a Context."""
    _check(lib().ned_h5_green_read(ctx.h, which, _b(path), _b(group)))


def write_context(ctx, path, group="", which=NED_G, nt=-1):
    """print_to_file of slices 0..nt of G or Sigma of a Context (nt < 0: all)."""
    _check(lib().ned_h5_green_write(ctx.h, which, _b(path), _b(group), nt))


__all__ = ["dims", "read_green", "write_green", "read_context", "write_context", "NED_G", "NED_SIGMA"]
