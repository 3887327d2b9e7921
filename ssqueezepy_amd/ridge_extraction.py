"""Ridge extraction from a time-frequency map (ridge_extraction.py:11-232):
forward-backward tracking of the minimum of -log normalized energy plus a
squared-jump penalty, through ``ssqk_ridge_track``.  The penalty grid is
the reference's host float32 arithmetic (log scales for a CWT)."""
import numpy as np

from . import _lib

__all__ = ['extract_ridges']

EPS32 = np.finfo(np.float32).eps


def extract_ridges(Tf, scales, penalty=2., n_ridges=1, bw=15, transform='cwt',
                   get_params=False, parallel=True):
    """Ridge indices [N, n_ridges] (and with ``get_params`` the scales or
    frequencies along them and their energies) of ``Tf`` [nf, N].

    The backward pass is the reference's serial one (``parallel=False``:
    the last frequency within eps of the accumulated penalty); its parallel
    pass lets any such frequency win, and agrees wherever one does."""
    import torch
    if torch.is_tensor(Tf):
        T = Tf.detach()
    else:
        T = torch.as_tensor(np.asarray(Tf))
    if T.dtype == torch.complex128:
        raise NotImplementedError("extract_ridges on complex128 is not on the GPU path")
    if T.ndim != 2:
        raise ValueError("`Tf` must be 2D [frequencies, times] (got %s)" % (tuple(T.shape),))
    dev = T.device if T.is_cuda else torch.device('cuda', torch.cuda.current_device())
    T = T.to(dev).to(torch.complex64).contiguous()
    nf, N = T.shape
    # ridge_extraction.py:112-130 (float32: eps32, scales and penalty cast)
    dtype = np.float32
    scales = np.asarray(scales.detach().cpu() if hasattr(scales, 'detach') else scales,
                        dtype=dtype)
    scales_orig = scales.copy()
    s = (np.log(scales) if transform == 'cwt' else scales).squeeze()
    if s.ndim != 1 or len(s) != nf:
        raise ValueError("`scales` must hold %d values (got shape %s)" % (nf, scales.shape))
    s_d = torch.as_tensor(np.ascontiguousarray(s, dtype=dtype), device=dev)
    ridge = torch.empty((n_ridges, N), dtype=torch.int32, device=dev)
    ridge_e = (torch.empty((n_ridges, N), dtype=torch.float32, device=dev)
               if get_params else None)
    _lib.ridge_track(T, s_d, np.float32(penalty), EPS32, n_ridges, bw, ridge, ridge_e)
    ridge_idxs = ridge.cpu().numpy().T.astype(int)
    if not get_params:
        return ridge_idxs
    ridge_f = np.stack([scales_orig[ridge_idxs[:, i]] for i in range(n_ridges)], axis=1)
    return ridge_idxs, ridge_f.reshape(N, n_ridges), ridge_e.cpu().numpy().T.copy()
