"""Phase transform + synchrosqueeze of a user time-frequency transform
(experimental.py:145-262): ``phase_transform`` and ``phase_ssqueeze`` on a
given Wx and dWx, through the same kernels as ``ssq_cwt`` / ``ssq_stft``
(phase: ``ssqk_phase``; bins and squeeze: ``ssqk_bins_from_*``,
``ssqk_squeeze``)."""
import numpy as np

from .ssqueezing import ssqueeze

__all__ = ['phase_transform', 'phase_ssqueeze']

EPS32 = float(np.finfo(np.float32).eps)
EPS64 = float(np.finfo(np.float64).eps)


def phase_transform(Wx, dWx=None, difftype='trig', difforder=4, gamma=None, fs=1.,
                    Sfs=None, rpadded=False, padtype='reflect', N=None, n1=None,
                    get_w=False, transform='cwt'):
    """(w, Wx, dWx, Sfs, gamma) of experimental.py:192-262: w by phase_cwt /
    phase_stft when ``get_w``; a missing dWx of a CWT is trigdiff of Wx."""
    from ._ssq_cwt import phase_cwt
    from ._ssq_stft import phase_stft, _make_Sfs
    if transform == 'stft' and dWx is None:
        raise NotImplementedError("`phase_transform` without `dWx` for "
                                  "STFT is not currently supported.")
    if rpadded and N is None:
        raise ValueError("`rpadded=True` requires `N`")
    if getattr(Wx, 'ndim', np.ndim(Wx)) > 2 and get_w:
        raise NotImplementedError("`get_w=True` unsupported with batched input.")
    if difftype != 'trig':
        raise ValueError("GPU computation only supports `difftype = 'trig'`")
    if gamma is None:
        is64 = str(getattr(Wx, 'dtype', '')).endswith('complex128')
        gamma = 10 * (EPS64 if is64 else EPS32)
    w = None
    if transform == 'cwt':
        if dWx is None:
            # experimental.py:201-209: dWx by trigdiff of the given Wx
            from .utils import trigdiff
            if N is None and not rpadded:
                N = Wx.shape[-1]
            if n1 is None:
                from ._common import p2up
                n1 = p2up(N)[1]
            dWx = trigdiff(Wx, fs, padtype, rpadded, N=N, n1=n1, transform='cwt')
        if get_w:
            w = phase_cwt(Wx, dWx, difftype, gamma)
    else:
        if Sfs is None:
            Sfs = _make_Sfs(Wx.shape[-2], fs)
        if get_w:
            w = phase_stft(Wx, dWx, Sfs, gamma)
    return w, Wx, dWx, Sfs, gamma


def phase_ssqueeze(Wx, dWx=None, ssq_freqs=None, scales=None, Sfs=None, fs=1., t=None,
                   squeezing='sum', maprange=None, wavelet=None, gamma=None,
                   was_padded=True, flipud=False, rpadded=False, padtype=None, N=None,
                   n1=None, difftype=None, difforder=None, get_w=False, get_dWx=False,
                   transform='cwt'):
    """``phase_transform`` then ``ssqueeze`` (experimental.py:145-189).
    Returns (Tx, Wx, ssq_freqs, scales, Sfs, w, dWx)."""
    w, Wx, dWx, Sfs, gamma = phase_transform(
        Wx, dWx, difftype or 'trig', difforder=difforder, gamma=gamma, rpadded=rpadded,
        padtype=padtype, N=N, n1=n1, get_w=get_w, fs=fs, Sfs=Sfs, transform=transform)
    if w is not None and not get_dWx:
        dWx = None
    if maprange is None:
        maprange = 'peak' if transform == 'cwt' else 'maximal'
    Tx, ssq_freqs = ssqueeze(Wx, w, ssq_freqs, scales, Sfs, fs=fs, t=t,
                             squeezing=squeezing, maprange=maprange, wavelet=wavelet,
                             gamma=gamma, was_padded=was_padded, flipud=flipud, dWx=dWx,
                             transform=transform)
    return Tx, Wx, ssq_freqs, scales, Sfs, w, dWx
