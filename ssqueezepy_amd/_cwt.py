"""Continuous wavelet transform on the GPU (``cwt`` of _cwt.py:12-320).

The host side keeps the reference's argument handling, padding parameters and
scale grid; the transform itself (pad, FFT, psih generation, band-limited
inverse FFTs and unpadding) is one call into ``ssqk_cwt_exec``.
"""
import os

import numpy as np

from . import _lib
from ._common import WARN, _process_fs_and_t, p2up
from ._scales import process_scales
from ._tensors import device_constant, signal_to_device, to_output
from .wavelets import Wavelet, KIND_GMW
from ._autograd import CwtFunction, differentiable_input, wants_grad

__all__ = ['cwt']


def _wavelet_kernel_params(wavelet):
    p = wavelet.fn.params64
    if wavelet.kind == KIND_GMW:
        return (p['gamma'], p['beta'])
    return (p['mu'], 0.)


def _nan_checks(x, nan_checks):
    if nan_checks is None:
        nan_checks = bool(isinstance(x, np.ndarray))
    if nan_checks:
        if not isinstance(x, np.ndarray):
            raise ValueError("`nan_checks=True` requires NumPy input.")
        elif np.isnan(x.max()) or np.isinf(x.max()) or np.isinf(x.min()):
            WARN("found NaN or inf values in `x`; will zero")
            x = np.where(np.isfinite(x), x, 0)
    return x


def _padding(N, padtype):
    if padtype is None:
        if N & (N - 1):
            raise NotImplementedError("padtype=None needs a power-of-two length "
                                      "on the GPU path (got N=%d)" % N)
        return N, 0, 'zero'
    if padtype not in _lib.PAD_CODES:
        raise ValueError("`padtype` must be one of %s (got %s)" %
                         (tuple(_lib.PAD_CODES), padtype))
    up, n1, _ = p2up(N)
    return up, n1, padtype


def autotune_env():
    return os.environ.get('SSQK_AUTOTUNE', '0') == '1'


def _maybe_autotune(device, N, L, n1, pad, wavelet, scales64, dt, padtype):
    """SSQK_AUTOTUNE=1: tune the plan of this grid once (default scale grid
    of ssq_cwt only; other grids keep the cost model's plan)."""
    from . import _autotune
    key = _lib.plan_key(device, N, L, n1, pad, wavelet.kind, _wavelet_kernel_params(wavelet),
                        scales64, dt)
    if _autotune.tuned_key_done(key) or key in _lib._tuned:
        return
    from ._ssq_cwt import ssq_cwt_grid
    g = ssq_cwt_grid(N, wavelet, padtype=padtype)
    if np.array_equal(g['scales64'], scales64) and g['dt'] == dt:
        _autotune.autotune(N, wavelet, padtype=padtype)


def run_cwt(x_dev, wavelet, scales64, dt, padtype, Wx, dWx=None, w=None,
            bins=None, bp=None, gamma=0., strict=True):
    """Launch the CWT plan for device signal(s) ``x_dev`` (float32 [B,] N)."""
    N = x_dev.shape[-1]
    L, n1, pad = _padding(N, padtype)
    if autotune_env():
        _maybe_autotune(x_dev.device, N, L, n1, pad, wavelet, scales64, dt, padtype)
    plan = _lib.cwt_plan(x_dev.device, N, L, n1, pad, wavelet.kind,
                         _wavelet_kernel_params(wavelet), scales64, dt)
    plan.exec(x_dev, Wx, dWx=dWx, w=w, bins=bins, bp=bp, gamma=gamma,
              strict=strict)
    return plan


def cwt(x, wavelet='gmw', scales='log-piecewise', fs=None, t=None, nv=32,
        l1_norm=True, derivative=False, padtype='reflect', rpadded=False,
        vectorized=True, astensor=True, cache_wavelet=None, order=0, average=None,
        nan_checks=None, patience=0):
    """Continuous wavelet transform of 1D/2D ``x`` via FFT convolution with the
    frequency-domain wavelet at the padded length (_cwt.py:12-320).

    Returns (Wx, scales) or (Wx, scales, dWx); Wx [na, N] or [B, na, N]
    complex64 on the GPU (NumPy with astensor=False)."""
    import torch
    if isinstance(order, (tuple, list, range)) or order > 0:
        raise NotImplementedError("higher-order GMW CWT is not on the GPU path")
    if not l1_norm:
        raise NotImplementedError("l1_norm=False is not on the GPU path")
    if rpadded:
        raise NotImplementedError("rpadded=True is not on the GPU path")
    if not hasattr(x, 'ndim'):
        raise TypeError("`x` must be a numpy array or torch Tensor "
                        "(got %s)" % type(x))
    elif x.ndim not in (1, 2):
        raise ValueError("`x` must be 1D or 2D (got x.ndim == %s)" % x.ndim)
    grad = wants_grad(x)
    if not grad:
        x = _nan_checks(x, nan_checks)
    if not isinstance(scales, str):
        nv = None
    N = x.shape[-1]
    dt, *_ = _process_fs_and_t(fs, t, N=N)
    wavelet = Wavelet._init_if_not_isinstance(wavelet)
    scales = process_scales(scales, N, wavelet, nv=nv)
    scales64 = np.asarray(scales, dtype=np.float64).ravel()
    na = len(scales64)

    def run(x_dev):
        shape = (na, N) if x_dev.ndim == 1 else (x_dev.shape[0], na, N)
        Wx = torch.empty(shape, dtype=torch.complex64, device=x_dev.device)
        dWx = torch.empty_like(Wx) if derivative else None
        plan = run_cwt(x_dev, wavelet, scales64, dt, padtype, Wx, dWx=dWx)
        return Wx, ((dWx,) if derivative else ()), plan

    if grad:
        x_dev = differentiable_input(x, wavelet.dtype)
        outs = CwtFunction.apply(x_dev, run)
        Wx, dWx = outs[0], (outs[1] if derivative else None)
    else:
        x_dev = signal_to_device(x, wavelet.dtype)
        Wx, extra, _ = run(x_dev)
        dWx = extra[0] if derivative else None
    scales_out = device_constant(scales64.astype(wavelet.dtype), torch.float32,
                                 x_dev.device, copy=True)
    Wx, scales_out, dWx = [to_output(g, astensor) for g in (Wx, scales_out, dWx)]
    return (Wx, scales_out, dWx) if derivative else (Wx, scales_out)
