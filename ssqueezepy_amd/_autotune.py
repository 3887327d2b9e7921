"""Autotuner for the CWT plan: times candidate plan options on the GPU and
keeps the fastest for the (length, wavelet, scales) it was run for.

The plan's own choices come from a cost model (FFT "bits" per output of the
band, overlap-save and four-step paths) and fixed launch shapes; what the
model cannot see -- how well classes overlap on the side streams, how many
scales an overlap-save workgroup should carry, where the band /
overlap-save crossover really is on this GPU -- is measured here:

* side_streams  0..3 streams beside the caller's for the scale classes;
* os_group      scales per overlap-save workgroup (1..8): fewer means more
                workgroups (better fill) but more segment FFTs;
* os_bias       factor on the overlap-save cost in the path choice.

Each candidate builds a plan and runs ssq_cwt's device work (CWT with the
squeeze rows, then the squeeze) on a seeded chirp of the same length,
``reps`` times after one warm-up, timed with HIP events on the current
stream.  The winner is registered (``_lib.set_tuned``) so later ssq_cwt /
cwt calls with the same grid build their plans with it.  Setting
``SSQK_AUTOTUNE=1`` runs this on the first ssq_cwt call of each grid.
"""
import itertools
import time

import numpy as np
import torch

from . import _lib
from ._common import EPS32
from ._cwt import _padding, _wavelet_kernel_params
from ._ssq_cwt import _device_const, ssq_cwt_grid
from .wavelets import Wavelet

__all__ = ['autotune']

DEFAULT_SPACE = dict(side_streams=(0, 2, 3), os_group=(4, 8), os_bias=(0.8, 1.0, 1.25))
_done = {}


def _signal(N):
    n = np.arange(N, dtype=np.float64)
    rng = np.random.default_rng(0)
    x = (np.cos(2 * np.pi * (0.02 * n + 0.09 * n**2 / N)) + 0.5 * np.cos(2 * np.pi * 0.3 * n)
         + 0.1 * rng.standard_normal(N))
    return x.astype(np.float32)


def autotune(N, wavelet='gmw', scales='log-piecewise', nv=None, fs=None, t=None,
             padtype='reflect', maprange='peak', flipud=True, batch=1, space=None,
             reps=3, verbose=False):
    """Tune the plan for length-N signals (``batch`` per call) and register
    the best options.  Returns dict(best=..., ms=..., default_ms=...,
    results=[(opts dict, ms), ...])."""
    dev = torch.device('cuda', torch.cuda.current_device())
    wavelet = Wavelet._init_if_not_isinstance(wavelet)
    g = ssq_cwt_grid(N, wavelet, scales, nv, fs, t, None, maprange, padtype, flipud)
    wav = g['wavelet']
    L, n1, pad = _padding(N, padtype)
    args = (dev, N, L, n1, pad, wav.kind, _wavelet_kernel_params(wav), g['scales64'], g['dt'])
    key = _lib.plan_key(*args)
    na = g['na']
    x = torch.as_tensor(np.tile(_signal(N), (batch, 1)) if batch > 1 else _signal(N), device=dev)
    shape = (na, N) if batch == 1 else (batch, na, N)
    Wx = torch.empty(shape, dtype=torch.complex64, device=dev)
    Tx = torch.empty_like(Wx)
    bins = torch.empty(shape, dtype=torch.int16, device=dev)
    cst = _device_const(g, dev)
    gamma = 10 * EPS32
    stream = torch.cuda.current_stream(dev)
    space = dict(DEFAULT_SPACE, **(space or {}))
    names = list(space)
    results = []
    best = (None, float('inf'), None)
    default_ms = None
    t0 = time.perf_counter()
    for combo in itertools.product(*(space[k] for k in names)):
        o = dict(zip(names, combo))
        opts = _lib.plan_opts(**o)
        plan = _lib.CwtPlan(*args, opts=opts)
        ws = torch.empty(plan.workspace_bytes(batch), dtype=torch.uint8, device=dev)

        def run():
            plan.exec(x, Wx, bins=bins, bp=g['bp'], gamma=gamma, strict=True, workspace=ws)
            _lib.squeeze(Wx, bins, cst, na, Tx)

        run()
        torch.cuda.synchronize(dev)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            run()
        e1.record(stream)
        torch.cuda.synchronize(dev)
        ms = e0.elapsed_time(e1) / reps
        results.append((o, ms))
        if o == dict(side_streams=3, os_group=8, os_bias=1.0):
            default_ms = ms
        if verbose:
            print("autotune N=%d %s: %.4f ms" % (N, o, ms), flush=True)
        if ms < best[1]:
            best = (o, ms, plan)
        del ws
    _lib.set_tuned(key, _lib.plan_opts(**best[0]), best[2])
    _done[key] = best[0]
    return dict(best=best[0], ms=best[1], default_ms=default_ms, results=results,
                seconds=time.perf_counter() - t0)


def tuned_key_done(key):
    return key in _done
