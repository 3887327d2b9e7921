"""Small host-side helpers shared by the transforms.

Reference counterparts: ssqueezepy/utils/common.py:1-158 (p2up, padsignal,
EPS constants, WARN/NOTE) and utils/cwt_utils.py:698-718 (_process_fs_and_t).
"""
import logging
import numpy as np

logging.basicConfig(format='')
_log = logging.getLogger('ssqueezepy_amd')

pi = np.pi
EPS32 = np.finfo(np.float32).eps
EPS64 = np.finfo(np.float64).eps


def WARN(msg):
    _log.warning("WARNING: %s" % msg)


def NOTE(msg):
    _log.warning("NOTE: %s" % msg)


def assert_is_one_of(x, name, supported, e=ValueError):
    if x not in supported:
        raise e("`{}` must be one of: {} (got {})".format(
            name, ', '.join(map(str, supported)), x))


def p2up(n):
    """Next power of two above ``n`` (rounded in log2) and the left/right pad
    lengths that centre ``n`` samples in it (utils/common.py:32-51)."""
    up = int(2**(1 + np.round(np.log2(n))))
    n2 = int((up - n) // 2)
    n1 = int(up - n - n2)
    return up, n1, n2


def pad_lengths(N, padlength=None):
    """(n_up, n1, n2) for ``padsignal`` (utils/common.py:100-111)."""
    if padlength is None:
        return p2up(N)
    n_up = int(padlength)
    if abs(padlength - N) % 2 == 0:
        n1 = n2 = (n_up - N) // 2
    else:
        n2 = (n_up - N) // 2
        n1 = n2 + 1
    return n_up, int(n1), int(n2)


def reflect_index(idx, N):
    """Index map of numpy's ``mode='reflect'`` padding for any pad length:
    the signal is extended with period ``2(N-1)`` mirrored about its end
    samples (edges not repeated)."""
    idx = np.asarray(idx, dtype=np.int64)
    if N == 1:
        return np.zeros_like(idx)
    P = 2 * (N - 1)
    m = np.mod(idx, P)
    return np.where(m < N, m, P - m)


def padsignal_np(x, padtype='reflect', padlength=None):
    """Host padding of a 1D/2D numpy signal (utils/common.py:54-158)."""
    N = x.shape[-1]
    n_up, n1, n2 = pad_lengths(N, padlength)
    width = (n1, n2) if x.ndim == 1 else [(0, 0), (n1, n2)]
    if padtype == 'zero':
        return np.pad(x, width), n_up, n1, n2
    if padtype == 'reflect':
        return np.pad(x, width, mode='reflect'), n_up, n1, n2
    if padtype == 'replicate':
        return np.pad(x, width, mode='edge'), n_up, n1, n2
    if padtype == 'wrap':
        return np.pad(x, width, mode='wrap'), n_up, n1, n2
    if padtype == 'symmetric':
        return np.pad(x, width, mode='symmetric'), n_up, n1, n2
    raise ValueError("unsupported `padtype` %s" % padtype)


def _process_fs_and_t(fs, t, N):
    """(dt, fs, t) from ``fs`` or a uniform time vector ``t``
    (utils/cwt_utils.py:698-718)."""
    if fs is not None and t is not None:
        WARN("`t` will override `fs` (both were passed)")
    if t is not None:
        t = np.asarray(t)
        if len(t) != N:
            raise Exception("`t` must be of same length as `x` "
                            "(%s != %s)" % (len(t), N))
        elif not np.mean(np.abs(np.diff(t, 2, axis=0))) < 1e-7:
            raise Exception("Time vector `t` must be uniformly sampled.")
        fs = 1 / (t[1] - t[0])
    else:
        if fs is None:
            fs = 1
        elif fs <= 0:
            raise ValueError("`fs` must be > 0")
    dt = 1 / fs
    return dt, fs, t
