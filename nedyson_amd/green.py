"""Host-side contour storage with the packing of the reference's green_func.

Mirrors src/greens.h:17-112 and src/greens.cpp:34-56, 177-195:

* ``ret``  lower triangle G^R(i, j<=i) at packed index i(i+1)/2 + j
* ``les``  upper triangle G^<(i<=j, j) at packed index j(j+1)/2 + i
* ``tv``   G^rceil(i, m), shape (nt+1, ntau+1, nao, nao)
* ``mat``  G^M(m), shape (ntau+1, nao, nao)

Every element is a row-major nao x nao complex128 block.  The arrays are plain
numpy; the device copy of the same layout lives in the C-ABI context.
"""
import numpy as np


class GreenFunc:
    def __init__(self, nt, ntau, nao, sig=-1):
        if nao <= 0 or nt < -1 or ntau < 0 or sig * sig != 1:
            raise ValueError("green_func: bad sizes (src/greens.cpp:35)")
        self.nt, self.ntau, self.nao, self.sig = nt, ntau, nao, sig
        ntri = (nt + 1) * (nt + 2) // 2 if nt >= 0 else 0
        self.ret = np.zeros((ntri, nao, nao), dtype=np.complex128)
        self.les = np.zeros((ntri, nao, nao), dtype=np.complex128)
        self.tv = np.zeros((max(nt + 1, 0), ntau + 1, nao, nao), dtype=np.complex128)
        self.mat = np.zeros((ntau + 1, nao, nao), dtype=np.complex128)

    @staticmethod
    def tri(i):
        return i * (i + 1) // 2

    def retptr(self, i, j):
        assert 0 <= j <= i <= self.nt
        return self.ret[i * (i + 1) // 2 + j]

    def lesptr(self, i, j):
        assert 0 <= i <= j <= self.nt
        return self.les[j * (j + 1) // 2 + i]

    def tvptr(self, i, m):
        return self.tv[i, m]

    def matptr(self, m):
        return self.mat[m]

    def ret_row(self, t):
        """G^R(t, 0..t): one contiguous (t+1, nao, nao) slab."""
        s = t * (t + 1) // 2
        return self.ret[s:s + t + 1]

    def les_col(self, t):
        """G^<(0..t, t): one contiguous (t+1, nao, nao) slab."""
        s = t * (t + 1) // 2
        return self.les[s:s + t + 1]

    def get_dm(self, i):
        """Density matrix, src/greens.h:56-65."""
        if i == -1:
            return -self.mat[self.ntau].copy()
        return 1j * self.sig * self.lesptr(i, i)

    def set_tstp_zero(self, tstp):
        """src/greens.h:76-87."""
        if tstp == -1:
            self.mat[:] = 0
        else:
            self.ret_row(tstp)[:] = 0
            self.les_col(tstp)[:] = 0
            self.tv[tstp] = 0

    def copy(self):
        g = GreenFunc.__new__(GreenFunc)
        g.nt, g.ntau, g.nao, g.sig = self.nt, self.ntau, self.nao, self.sig
        g.ret, g.les, g.tv, g.mat = self.ret.copy(), self.les.copy(), self.tv.copy(), self.mat.copy()
        return g
