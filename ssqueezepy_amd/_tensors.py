"""Host <-> device conversions shared by the transforms.

Inputs may be NumPy arrays or torch tensors (any device); the transforms run
on the current CUDA (HIP) device and return torch tensors when
``astensor=True``, NumPy arrays otherwise, as the reference does under
SSQ_GPU=1 (ssqueezepy/utils/backend.py:26-42, _ssq_cwt.py:297-300).
"""
from collections import OrderedDict
import threading

import numpy as np

_consts = OrderedDict()
_consts_lock = threading.Lock()


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("ssqueezepy_amd needs a ROCm GPU (torch.cuda is not "
                           "available)")
    return torch


def device_of(x):
    torch = _torch()
    if isinstance(x, torch.Tensor) and x.is_cuda:
        return x.device
    return torch.device('cuda', torch.cuda.current_device())


def as_device_tensor(x, like=None, complex_=False, real=False, device=None):
    """Contiguous complex64 / float32 device tensor of ``x``."""
    torch = _torch()
    dev = device or (like.device if like is not None else device_of(x))
    if isinstance(x, np.ndarray):
        t = torch.from_numpy(np.ascontiguousarray(x))
    elif isinstance(x, torch.Tensor):
        t = x.detach()
    else:
        raise TypeError("expected a NumPy array or torch Tensor (got %s)" % type(x))
    if complex_:
        if t.dtype == torch.complex128:
            raise NotImplementedError("complex128 inputs: the GPU path computes "
                                      "in complex64")
        t = t.to(device=dev, dtype=torch.complex64)
    elif real:
        if t.dtype == torch.float64:
            raise NotImplementedError("float64 inputs: the GPU path computes in "
                                      "float32")
        t = t.to(device=dev, dtype=torch.float32)
    else:
        t = t.to(device=dev)
    return t.contiguous()


def signal_to_device(x, dtype='float32'):
    """Signal ``x`` (1D/2D) as a contiguous float32 device tensor, cast to the
    transform dtype like ``S.asarray(x, dtype)`` (_cwt.py:255-258)."""
    torch = _torch()
    if dtype != 'float32':
        raise NotImplementedError("only float32 transforms run on the GPU path "
                                  "(wavelet/STFT dtype=%s)" % dtype)
    if isinstance(x, np.ndarray):
        t = torch.from_numpy(np.ascontiguousarray(x.astype(np.float32)))
        t = t.to(device_of(x))
    elif isinstance(x, torch.Tensor):
        t = x.detach().to(device=device_of(x), dtype=torch.float32)
    else:
        raise TypeError("`x` must be a numpy array or torch Tensor "
                        "(got %s)" % type(x))
    return t.contiguous()


def to_output(t, astensor):
    if t is None or astensor:
        return t
    torch = _torch()
    if isinstance(t, torch.Tensor):
        return t.cpu().numpy()
    return t


def device_constant(arr, dtype, device, copy=False):
    """Small host array ``arr`` as a ``dtype`` tensor on ``device``.

    A pageable host->device copy makes the host wait for the stream, which
    keeps it from queueing the next transform while the GPU works; these
    arrays (scales, squeeze weights, windows, frequency grids) repeat from
    call to call, so each is uploaded once and kept (LRU of 64).  With
    ``copy`` the caller gets its own device copy (an asynchronous
    device-to-device copy), for arrays handed back to the user."""
    torch = _torch()
    a = np.ascontiguousarray(np.asarray(arr))
    key = (str(device), str(dtype), a.dtype.str, a.shape, a.tobytes())
    with _consts_lock:
        t = _consts.get(key)
        if t is not None:
            _consts.move_to_end(key)
    if t is None:
        t = torch.as_tensor(a, device=device).to(dtype)
        with _consts_lock:
            _consts[key] = t
            while len(_consts) > 64:
                _consts.popitem(last=False)
    return t.clone() if copy else t
