"""Utilities of the reference's ``ssqueezepy.utils`` that sit on the hot
path's outputs: ``trigdiff`` (utils/common.py:161-245, on the GPU through
``ssqk_trigdiff``), ``p2up`` / ``padsignal`` and the admissibility
constants."""
import numpy as np

from . import _lib
from ._common import p2up, padsignal_np as padsignal
from ._inverse import adm_ssq, adm_cwt, integrate_analytic

__all__ = ['trigdiff', 'p2up', 'padsignal', 'adm_ssq', 'adm_cwt', 'integrate_analytic']


def trigdiff(A, fs=1., padtype=None, rpadded=None, N=None, n1=None, window=None,
             transform='cwt'):
    """Frequency-domain derivative of the rows of ``A`` (2D, or 3D batched):
    pad (``padtype``, default 'reflect' unless ``rpadded``), FFT, times
    i xi fs, inverse FFT, unpad to ``[n1, n1 + N)`` (utils/common.py:
    161-245).  Padded lengths up to 8192 on the GPU path."""
    import torch
    if transform == 'stft':
        raise NotImplementedError("`transform='stft'` is currently not supported.")
    T = A.detach() if torch.is_tensor(A) else torch.as_tensor(np.asarray(A))
    assert T.ndim in (2, 3), T.shape
    if rpadded and N is None:
        raise ValueError("must pass `N` if `rpadded`")
    rpadded = rpadded or False
    padtype = padtype or ('reflect' if not rpadded else None)
    if T.dtype == torch.complex128:
        raise NotImplementedError("trigdiff on complex128 is not on the GPU path")
    dev = T.device if T.is_cuda else torch.device('cuda', torch.cuda.current_device())
    T = T.to(dev).to(torch.complex64).contiguous()
    # (the reference pads with padsignal, whose n1 replaces the given one,
    # defaults N to the padded length, and unpads as A_diff[..., n1:n1 + N],
    # which a slice clips at L)
    if padtype is not None:
        L, n1, _ = p2up(T.shape[-1])
        Nreq = N if N is not None else L
    else:
        L = T.shape[-1]
        Nreq = N if N is not None else L
        if n1 is None:
            n1 = p2up(Nreq)[1]
    Nout = max(0, min(Nreq, L - n1))
    out = torch.empty(T.shape[:-1] + (Nout,), dtype=torch.complex64, device=dev)
    if Nout == 0:
        return out
    _lib.trigdiff(T, Nout, L, n1, padtype, fs, out)
    return out
