"""Timed CPU baseline: the reference's parallel CPU pipeline restated --
BENCHMARK/TEST INFRASTRUCTURE ONLY.

Mirrors ssq_cwt with SSQ_PARALLEL=1 on the CPU (_ssq_cwt.py:190-310,
_cwt.py:167-177): scipy.fft with workers = all host threads (the reference's
FFT backend and threading, utils/fft_utils.py:135-140,164,190), psih in the
wavelet's float32 dtype with NumPy, and the column-parallel squeeze loop
(algos.py:879-984) as the OpenMP C port in cpu_ssq.c.
"""
import ctypes
import os
import subprocess

import numpy as np
import scipy.fft as sfft

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, '_build', 'libcpu_ssq.so')
_lib = None


def build():
    subprocess.run(['make', '-s', '-C', HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        f = lib.ssq_squeeze_par
        P = ctypes.c_void_p
        f.argtypes = [P, P, P, P, P, ctypes.c_int64, ctypes.c_int64, ctypes.c_double,
                      ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                      ctypes.c_double, ctypes.c_int64, ctypes.c_int64, ctypes.c_int]
        f.restype = None
        _lib = lib
    return _lib


def squeeze_par(Wx, dWx, const, params, omax, flipud, gamma, Sfs=None):
    """Tx via the C port; Wx, dWx complex64 [na, N]; params from
    ssq_oracle.bin_params."""
    Wx = np.ascontiguousarray(Wx, dtype=np.complex64)
    dWx = np.ascontiguousarray(dWx, dtype=np.complex64)
    na, N = Wx.shape
    out = np.zeros((na, N), dtype=np.complex64)
    cst = np.ascontiguousarray(np.broadcast_to(np.asarray(const, np.float32).ravel(),
                                               (na,)))
    kind = {'lin': 0, 'log': 1, 'piecewise': 2}[params['kind']]
    if kind == 0:
        v0, d0, v1, d1, idx1 = params['vmin'], params['dv'], 0., 1., 0
    elif kind == 1:
        v0, d0, v1, d1, idx1 = params['vlmin'], params['dvl'], 0., 1., 0
    else:
        v0, d0 = params['vlmin0'], params['dvl0']
        v1, d1, idx1 = params['vlmin1'], params['dvl1'], params['idx1']
    sf = None if Sfs is None else np.ascontiguousarray(Sfs, dtype=np.float32)
    load().ssq_squeeze_par(Wx.ctypes.data, dWx.ctypes.data,
                           None if sf is None else sf.ctypes.data, out.ctypes.data,
                           cst.ctypes.data, na, N, float(gamma), kind, v0, d0, v1, d1,
                           int(idx1), int(omax), int(bool(flipud)))
    return out


def gmw_psih32(w, gamma=3., beta=60.):
    """_gmw.py:203-210 in float32 (the reference's default dtype)."""
    w = np.asarray(w, dtype=np.float32)
    wc = (beta / gamma)**(1 / gamma)
    g, b = np.float32(gamma), np.float32(beta)
    wcf, wcl = np.float32(wc), np.float32(np.log(wc))
    nonneg = w >= 0
    w = w * nonneg
    with np.errstate(divide='ignore'):
        return 2 * np.exp(-b * wcl + wcf**g + b * np.log(w) - w**g) * nonneg


class SsqCwtCPU:
    """Reference CPU pipeline for one signal length and scale grid (float32)."""

    def __init__(self, N, scales, const, params, omax, flipud=True, gamma=None,
                 psih32=gmw_psih32, workers=None):
        self.N = N
        up = int(2**(1 + np.round(np.log2(N))))
        n2 = (up - N) // 2
        self.L, self.n1, self.n2 = up, up - N - n2, n2
        self.workers = workers or os.cpu_count()
        xi = np.arange(up, dtype=np.float64)
        xi[up // 2 + 1:] -= up
        self.xi = (xi * (2 * np.pi / up)).astype(np.float32)
        s = np.asarray(scales, dtype=np.float32).reshape(-1, 1)
        P = psih32(s * self.xi)
        P[:, up // 2] /= 2
        self.Psih = P                      # cached, as cache_wavelet=True
        self.const, self.params, self.omax = const, params, omax
        self.flipud = flipud
        self.gamma = gamma if gamma is not None else 10 * np.finfo(np.float32).eps

    def __call__(self, x):
        xp = np.pad(np.asarray(x, np.float32), (self.n1, self.n2), mode='reflect')
        xh = sfft.fft(xp, workers=self.workers)
        Psih_xh = self.Psih * xh
        Wx = sfft.ifft(Psih_xh, workers=self.workers)[:, self.n1:self.n1 + self.N]
        Psih_xh *= (1j * self.xi)
        dWx = sfft.ifft(Psih_xh, workers=self.workers)[:, self.n1:self.n1 + self.N]
        Wx = np.ascontiguousarray(Wx)
        Tx = squeeze_par(Wx, dWx, self.const, self.params, self.omax, self.flipud,
                         self.gamma)
        return Tx, Wx
