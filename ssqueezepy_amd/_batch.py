"""ssq_cwt over many signals of uneven length in one queued pass.

The reference takes a batch only as a 2D array of equal-length signals
(_ssq_cwt.py:12-310 broadcasts over the leading axis); signals of different
lengths need one call each, and at N ~ 4k-64k each call is a few dozen small
kernels that leave most of the GPU idle.  ``ssq_cwt_batch`` keeps every
signal's own grid (scales and ssq_freqs from its length, as the reference
computes them) and plan, preallocates all inputs, outputs and workspaces,
and queues the signals' transforms round robin on several streams so that
their kernels run concurrently.  Runners are cached per (lengths, wavelet,
options); a new call with the same lengths copies the signals into the
runner's input buffers and queues the pass again.  Outputs of a pass
overwrite the previous pass's tensors unless ``copy=True``.

``graph=True`` (capturing the pass into a HIP graph) is switched off: its
first run ended the process with a segmentation fault inside the runtime's
end-of-capture call, cause not yet found, so it raises NotImplementedError.
"""
from collections import OrderedDict

import numpy as np
import torch

from . import _lib
from ._common import EPS32
from ._cwt import _padding, _wavelet_kernel_params
from ._ssq_cwt import _device_const, ssq_cwt_grid
from ._tensors import device_constant, to_output
from .wavelets import Wavelet

__all__ = ['ssq_cwt_batch']

_runners = OrderedDict()


class _BatchRunner:
    def __init__(self, lengths, wavelet, scales, nv, fs, t, maprange, padtype, flipud,
                 gamma, nstreams, device):
        self.device = device
        self.items = []
        for N in lengths:
            g = ssq_cwt_grid(N, wavelet, scales, nv, fs, t, None, maprange, padtype,
                             flipud)
            wav = g['wavelet']
            L, n1, pad = _padding(N, padtype)
            plan = _lib.cwt_plan(device, N, L, n1, pad, wav.kind,
                                 _wavelet_kernel_params(wav), g['scales64'], g['dt'])
            na = g['na']
            it = dict(
                N=N, g=g, plan=plan, na=na,
                x=torch.zeros(N, dtype=torch.float32, device=device),
                Wx=torch.empty((na, N), dtype=torch.complex64, device=device),
                Tx=torch.empty((na, N), dtype=torch.complex64, device=device),
                bins=torch.empty((na, N), dtype=torch.int16, device=device),
                ws=torch.empty(plan.workspace_bytes(1), dtype=torch.uint8, device=device),
                cst=_device_const(g, device),
                scales=device_constant(g['scales'], torch.float32, device),
                freqs=g['ssq_freqs'][::-1])
            self.items.append(it)
        self.gamma = gamma
        self.streams = [torch.cuda.Stream(device=device) for _ in range(nstreams)]

    def _launch(self):
        """Queue every signal's transform and squeeze; the streams fork from
        and join into the current stream."""
        cur = torch.cuda.current_stream(self.device)
        for s in self.streams:
            s.wait_stream(cur)
        for k, it in enumerate(self.items):
            with torch.cuda.stream(self.streams[k % len(self.streams)]):
                it['plan'].exec(it['x'], it['Wx'], bins=it['bins'], bp=it['g']['bp'],
                                gamma=self.gamma, strict=True, workspace=it['ws'])
                _lib.squeeze(it['Wx'], it['bins'], it['cst'], it['na'], it['Tx'])
        for s in self.streams:
            cur.wait_stream(s)

    def run(self, xs):
        for it, x in zip(self.items, xs):
            it['x'].copy_(x, non_blocking=True)
        self._launch()


def _as_signal(x, device):
    if isinstance(x, np.ndarray):
        return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(device)
    if isinstance(x, torch.Tensor):
        return x.detach().to(device=device, dtype=torch.float32).contiguous()
    raise TypeError("signals must be NumPy arrays or torch tensors (got %s)" % type(x))


def ssq_cwt_batch(xs, wavelet='gmw', scales='log-piecewise', nv=None, fs=None, t=None,
                  padtype='reflect', maprange='peak', flipud=True, gamma=None,
                  astensor=True, graph=False, streams=4, copy=False):
    """Synchrosqueezed CWT of each 1D signal in ``xs`` (any lengths).

    Returns a list of (Tx, Wx, ssq_freqs, scales), one per signal, equal to
    ``ssq_cwt(x, wavelet, scales, nv, fs, t, padtype=padtype, maprange=maprange,
    flipud=flipud, gamma=gamma)`` of each signal (same plans and kernels).
    """
    if not torch.cuda.is_available():
        raise RuntimeError("ssqueezepy_amd needs a ROCm GPU")
    if graph:
        raise NotImplementedError("graph capture of the batched pass is switched off "
                                  "(end of capture crashed; cause not found yet)")
    xs = list(xs)
    if any(getattr(x, 'ndim', 1) != 1 for x in xs):
        raise ValueError("ssq_cwt_batch takes a sequence of 1D signals")
    device = torch.device('cuda', torch.cuda.current_device())
    wavelet = Wavelet._init_if_not_isinstance(wavelet)
    gamma = 10 * EPS32 if gamma is None else gamma
    lengths = tuple(int(x.shape[-1]) for x in xs)
    key = (str(device), lengths, wavelet.kind, tuple(_wavelet_kernel_params(wavelet)),
           wavelet.dtype, str(scales), nv, fs,
           None if t is None else np.asarray(t).tobytes(), padtype, maprange, flipud,
           float(gamma), int(streams))
    runner = _runners.get(key)
    if runner is None:
        runner = _BatchRunner(lengths, wavelet, scales, nv, fs, t, maprange, padtype,
                              flipud, gamma, int(streams), device)
        _runners[key] = runner
        while len(_runners) > 4:
            _runners.popitem(last=False)
    else:
        _runners.move_to_end(key)
    runner.run([_as_signal(x, device) for x in xs])
    out = []
    for it in runner.items:
        Tx, Wx = it['Tx'], it['Wx']
        sc = it['scales'].clone()
        if copy:
            Tx, Wx = Tx.clone(), Wx.clone()
        out.append((to_output(Tx, astensor), to_output(Wx, astensor), it['freqs'],
                    to_output(sc, astensor) if astensor else it['g']['scales'].copy()))
    return out
