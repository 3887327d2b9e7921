"""GPU counterparts of the reference's squeeze kernels (algos.py:126-250,
706-856): ``ssqueeze_fast``, ``indexed_sum_onfly``, ``phase_cwt_gpu`` and
``phase_stft_gpu`` with the reference signatures."""
import numpy as np

from . import _lib
from ._tensors import as_device_tensor, device_constant
from .ssqueezing import bin_params_from_freqs

__all__ = ['ssqueeze_fast', 'indexed_sum_onfly', 'phase_cwt_gpu',
           'phase_stft_gpu']


def _const_tensor(const, na, like):
    import torch
    c = np.asarray(const.detach().cpu() if hasattr(const, 'detach') else const,
                   dtype=np.float64).ravel()
    if c.size != na:
        c = np.full(na, float(c[0]) if c.size else 1.)
    return device_constant(c, torch.float32, like.device)


def ssqueeze_fast(Wx, dWx, ssq_freqs, const, logscale=False, flipud=False,
                  gamma=None, out=None, Sfs=None, parallel=None):
    """Phase transform, bin search and ordered sum in two passes
    (algos.py:126-150).  Returns Tx (complex64, device)."""
    import torch
    Wx = as_device_tensor(Wx, complex_=True)
    dWx = as_device_tensor(dWx, like=Wx, complex_=True)
    na = Wx.shape[-2]
    bp = bin_params_from_freqs(np.asarray(ssq_freqs), logscale, flipud, na - 1)
    bins = torch.empty(Wx.shape, dtype=torch.int16, device=Wx.device)
    Sfs_t = None
    if Sfs is not None:
        Sfs_t = as_device_tensor(np.asarray(Sfs.detach().cpu() if hasattr(
            Sfs, 'detach') else Sfs, dtype=np.float32), like=Wx, real=True)
    _lib.bins_from_dwx(Wx, dWx, Sfs_t, bp, gamma, bins)
    Tx = out if out is not None else torch.empty_like(Wx)
    _lib.squeeze(Wx, bins, _const_tensor(const, na, Wx), na, Tx)
    return Tx


def indexed_sum_onfly(Wx, w, ssq_freqs, const=1, logscale=False, flipud=False,
                      out=None, parallel=None):
    """Bin search of a given phase transform and ordered sum
    (algos.py:153-169); points with inf ``w`` are skipped."""
    import torch
    Wx = as_device_tensor(Wx, complex_=True)
    w = as_device_tensor(w, like=Wx, real=True)
    na = Wx.shape[-2]
    bp = bin_params_from_freqs(np.asarray(ssq_freqs), logscale, flipud, na - 1)
    bins = torch.empty(Wx.shape, dtype=torch.int16, device=Wx.device)
    _lib.bins_from_w(w, bp, bins)
    Tx = out if out is not None else torch.empty_like(Wx)
    _lib.squeeze(Wx, bins, _const_tensor(const, na, Wx), na, Tx)
    return Tx


def phase_cwt_gpu(Wx, dWx, gamma):
    """|Im(dWx/Wx)| / 2pi, inf where |Wx| < gamma (algos.py:743-781)."""
    import torch
    Wx = as_device_tensor(Wx, complex_=True)
    dWx = as_device_tensor(dWx, like=Wx, complex_=True)
    w = torch.empty(Wx.shape, dtype=torch.float32, device=Wx.device)
    _lib.phase(Wx, dWx, None, gamma, w)
    return w


def phase_stft_gpu(Wx, dWx, Sfs, gamma):
    """|Sfs[i] - Im(dWx/Wx)/2pi|, inf where |Wx| < gamma (algos.py:818-856)."""
    import torch
    Wx = as_device_tensor(Wx, complex_=True)
    dWx = as_device_tensor(dWx, like=Wx, complex_=True)
    Sfs = as_device_tensor(np.asarray(Sfs.detach().cpu() if hasattr(Sfs, 'detach')
                                      else Sfs, dtype=np.float32), like=Wx, real=True)
    w = torch.empty(Wx.shape, dtype=torch.float32, device=Wx.device)
    _lib.phase(Wx, dWx, Sfs, gamma, w)
    return w
