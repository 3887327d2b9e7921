"""CWT scale grids (host, NumPy).

Restates ssqueezepy/utils/cwt_utils.py:66-580 (cwt_scalebounds,
process_scales, infer_scaletype, make_scales, logscale_transition_idx,
nv_from_scales, find_min_scale, find_max_scale, find_downsampling_scale) and
:630-695 (find_max_scale_alt).  The grids are API outputs and set the
synchrosqueezing bin edges, so every search keeps the reference's sampling,
dtype and tie-breaking; the committed golden fixtures pin the results.
"""
import numpy as np
from .configs import gdefaults
from ._common import WARN, assert_is_one_of, p2up, pi
from .wavelets import (Wavelet, center_frequency, find_maximum,
                       find_first_occurrence)

__all__ = ['cwt_scalebounds', 'process_scales', 'infer_scaletype',
           'make_scales', 'logscale_transition_idx', 'nv_from_scales',
           'find_min_scale', 'find_max_scale', 'find_downsampling_scale',
           'find_max_scale_alt']


def _asnumpy(x):
    if hasattr(x, 'detach'):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def cwt_scalebounds(wavelet, N, preset=None, min_cutoff=None, max_cutoff=None,
                    cutoff=None, bin_loc=None, bin_amp=None, use_padded_N=True,
                    viz=False):
    """(min_scale, max_scale) for which ``wavelet`` is well resolved
    (cwt_utils.py:66-188)."""
    defaults = dict(min_cutoff=.6, max_cutoff=.8, cutoff=-.5)
    if preset is not None:
        if any((min_cutoff, max_cutoff, cutoff)):
            WARN("`preset` will override `min_cutoff, max_cutoff, cutoff`")
        elif preset == 'minimal' and any((bin_amp, bin_loc)):
            WARN("`preset='minimal'` ignores `bin_amp` & `bin_loc`")
        assert_is_one_of(preset, 'preset', ('maximal', 'minimal', 'naive'))
        if preset in ('naive', 'maximal'):
            min_cutoff, max_cutoff = None, None
            if preset == 'maximal':
                cutoff = -.5
        else:
            min_cutoff, max_cutoff, cutoff = defaults.values()
    else:
        if min_cutoff is None:
            min_cutoff = defaults['min_cutoff']
        elif min_cutoff <= 0:
            raise ValueError("`min_cutoff` must be >0 (got %s)" % min_cutoff)
        if max_cutoff is None:
            max_cutoff = defaults['max_cutoff']
        elif max_cutoff < min_cutoff:
            raise ValueError("must have `max_cutoff > min_cutoff` "
                             "(got %s, %s)" % (max_cutoff, min_cutoff))
    bin_loc = bin_loc or (2 if preset == 'maximal' else None)
    bin_amp = bin_amp or (1 if preset == 'maximal' else None)
    cutoff = cutoff if (cutoff is not None) else defaults['cutoff']

    if preset == 'naive':
        return 1, N
    M = p2up(N)[0] if use_padded_N else N
    min_scale = find_min_scale(wavelet, cutoff=cutoff)
    if preset in ('minimal', None):
        max_scale = find_max_scale_alt(wavelet, M, min_cutoff=min_cutoff,
                                       max_cutoff=max_cutoff)
    else:
        max_scale = find_max_scale(wavelet, M, bin_loc=bin_loc,
                                   bin_amp=bin_amp)
    return min_scale, max_scale


def _assert_positive_integer(g, name=''):
    if not (g > 0 and float(g).is_integer()):
        raise ValueError(f"'{name}' must be a positive integer (got {g})")


def process_scales(scales, N, wavelet=None, nv=None, get_params=False,
                   use_padded_N=True):
    """Make scales from a string spec or validate an array; optionally return
    (scales, scaletype, na, nv) (cwt_utils.py:196-261)."""
    preset = None
    if isinstance(scales, str):
        if ':' in scales:
            scales, preset = scales.split(':')
        elif scales == 'log-piecewise':
            preset = 'maximal'
        assert_is_one_of(scales, 'scales', ('log', 'log-piecewise', 'linear'))
        if nv is None:
            nv = 32
        if wavelet is None:
            raise ValueError("must set `wavelet` if `scales` isn't array")
        scaletype = scales
    elif isinstance(scales, np.ndarray) or hasattr(scales, 'detach'):
        scales = _asnumpy(scales)
        if scales.squeeze().ndim != 1:
            raise ValueError("`scales`, if array, must be 1D "
                             "(got shape %s)" % str(scales.shape))
        scaletype, _nv = infer_scaletype(scales)
        if scaletype == 'log':
            if nv is not None and _nv != nv:
                raise Exception("`nv` used in `scales` differs from "
                                "`nv` passed (%s != %s)" % (_nv, nv))
            nv = _nv
        elif scaletype == 'log-piecewise':
            nv = _nv
        scales = scales.reshape(-1, 1)
    else:
        raise TypeError("`scales` must be a string or Numpy array "
                        "(got %s)" % type(scales))
    if nv is not None and not isinstance(nv, np.ndarray):
        _assert_positive_integer(nv, 'nv')
        nv = int(nv)

    if isinstance(scales, np.ndarray):
        scales = scales.reshape(-1, 1)
        return (scales if not get_params else
                (scales, scaletype, len(scales), nv))

    wavelet = Wavelet._init_if_not_isinstance(wavelet)
    min_scale, max_scale = cwt_scalebounds(wavelet, N=N, preset=preset,
                                           use_padded_N=use_padded_N)
    scales = make_scales(N, min_scale, max_scale, nv=nv, scaletype=scaletype,
                         wavelet=wavelet)
    na = len(scales)
    return (scales if not get_params else (scales, scaletype, na, nv))


def infer_scaletype(scales):
    """'log' (with nv), 'linear' or 'log-piecewise' (nv array)
    (cwt_utils.py:264-298)."""
    scales = _asnumpy(scales).reshape(-1, 1)
    if scales.dtype not in (np.float32, np.float64):
        raise TypeError("`scales.dtype` must be np.float32 or np.float64 "
                        "(got %s)" % scales.dtype)
    th_log = 4e-15 if scales.dtype == np.float64 else 8e-7
    th_lin = th_log * 1e3
    if np.mean(np.abs(np.diff(np.log(scales), 2, axis=0))) < th_log:
        scaletype = 'log'
        nv = int(np.round(1 / np.diff(np.log2(scales), axis=0)[0].squeeze()))
    elif np.mean(np.abs(np.diff(scales, 2, axis=0))) < th_lin:
        scaletype = 'linear'
        nv = None
    elif logscale_transition_idx(scales) is None:
        raise ValueError("could not infer `scaletype` from `scales`; "
                         "`scales` array must be linear or exponential.")
    else:
        scaletype = 'log-piecewise'
        nv = nv_from_scales(scales)
    return scaletype, nv


def make_scales(N, min_scale=None, max_scale=None, nv=32, scaletype='log',
                wavelet=None, downsample=None):
    """2^(k/nv) grid between min and max scale; 'log-piecewise' keeps
    scales[:idx] then every ``downsample``-th (cwt_utils.py:301-372)."""
    if scaletype == 'log-piecewise' and wavelet is None:
        raise ValueError("must pass `wavelet` for `scaletype == 'log-piecewise'`")
    if min_scale is None and max_scale is None and wavelet is not None:
        min_scale, max_scale = cwt_scalebounds(wavelet, N, use_padded_N=True)
    else:
        min_scale = min_scale or 1
        max_scale = max_scale or N
    downsample = int(gdefaults('utils.cwt_utils.make_scales',
                               downsample=downsample)['downsample'])
    na = int(np.ceil(nv * np.log2(max_scale / min_scale)))
    mn_pow = int(np.floor(nv * np.log2(min_scale)))
    mx_pow = mn_pow + na
    if scaletype == 'log':
        scales = 2 ** (np.arange(mn_pow, mx_pow) / nv)
    elif scaletype == 'log-piecewise':
        scales = 2 ** (np.arange(mn_pow, mx_pow) / nv)
        idx = find_downsampling_scale(wavelet, scales)
        if idx is not None:
            scales1 = scales[:idx]
            scales2 = scales[idx + downsample - 1::downsample]
            scales = np.hstack([scales1, scales2])
    elif scaletype == 'linear':
        min_scale, max_scale = 2**(mn_pow / nv), 2**(mx_pow / nv)
        na = int(np.ceil(max_scale / min_scale))
        scales = np.linspace(min_scale, max_scale, na)
    else:
        raise ValueError("`scaletype` must be 'log' or 'linear'; "
                         "got: %s" % scaletype)
    return scales.reshape(-1, 1)


def logscale_transition_idx(scales):
    """Index splitting a two-segment log grid (cwt_utils.py:375-394)."""
    scales = _asnumpy(scales)
    scales_diff2 = np.abs(np.diff(np.log(scales), 2, axis=0))
    idx = np.argmax(scales_diff2) + 2
    diff2_max = scales_diff2.max()
    scales_diff2[idx - 2] = 0
    th = 1e-14 if scales.dtype == np.float64 else 1e-6
    if not np.any(diff2_max > 100 * np.abs(scales_diff2).mean()):
        return None
    elif not np.all(np.abs(scales_diff2) < th):
        return None
    return idx


def nv_from_scales(scales):
    """Voices per octave of each scale, (na, 1) (cwt_utils.py:397-409)."""
    scales = _asnumpy(scales)
    logdiffs = 1 / np.diff(np.log2(scales), axis=0)
    nv = np.vstack([logdiffs[:1], logdiffs])
    idx = logscale_transition_idx(scales)
    if idx is not None:
        nv_transition_idx = np.argmax(np.abs(np.diff(nv, axis=0))) + 1
        assert nv_transition_idx == idx, "%s != %s" % (nv_transition_idx, idx)
    return nv


def find_min_scale(wavelet, cutoff=1):
    """Scale whose Nyquist sample sits at ``cutoff`` of the peak
    (cwt_utils.py:412-431)."""
    wavelet = Wavelet._init_if_not_isinstance(wavelet)
    w_peak, peak = find_maximum(wavelet.fn)
    if cutoff > 0:
        step_start, step_limit = w_peak, 10 * w_peak
    else:
        step_start, step_limit = 0, w_peak
    w_cutoff, _ = find_first_occurrence(wavelet.fn, value=abs(cutoff) * peak,
                                        step_start=step_start,
                                        step_limit=step_limit)
    return w_cutoff / pi


def find_max_scale(wavelet, N, bin_loc=1, bin_amp=1):
    """Scale at which psih reaches ``bin_amp`` of its max at bin ``bin_loc``
    (cwt_utils.py:434-456)."""
    wavelet = Wavelet._init_if_not_isinstance(wavelet)
    wc_ct = center_frequency(wavelet, kind='peak-ct', N=N)
    scalec_ct = (4 / pi) * wc_ct
    psih = np.asarray(wavelet(scale=scalec_ct, N=N)[:N // 2 + 1])
    xi = np.asarray(wavelet.xifn(scalec_ct, N))
    midx = np.argmax(psih)
    w_bin = xi[np.where(psih[:midx] < psih.max() * bin_amp)[0][-1]]
    return scalec_ct * (w_bin / xi[bin_loc])


def find_downsampling_scale(wavelet, scales, span=5, tol=3, method='sum',
                            nonzero_th=.02, nonzero_tol=4., N=None):
    """Index past which adjacent wavelets are too redundant
    (cwt_utils.py:459-580), evaluated at N=2048 by default."""
    def check_group(psihs_peaks, joint_peak, method, tol):
        distances = np.abs(psihs_peaks[1] - joint_peak)
        if method == 'any':
            return distances.max() < tol
        elif method == 'all':
            return not np.all(distances > tol)
        return distances.sum() < tol

    assert_is_one_of(method, 'method', ('any', 'all', 'sum'))
    if isinstance(wavelet, np.ndarray):
        Psih = wavelet
    else:
        wavelet = Wavelet._init_if_not_isinstance(wavelet)
        Psih = wavelet(scale=scales, N=N or 2048)
    Psih = np.asarray(Psih)
    if len(Psih) != len(scales):
        raise ValueError("len(Psih) != len(scales) "
                         "(%s != %s)" % (len(Psih), len(scales)))
    Psih = Psih[:, :Psih.shape[1] // 2]
    n_scales = len(Psih)
    n_groups = n_scales - span - 1
    for i in range(n_groups):
        psihs = Psih[i:i + span]
        psihs_nonzeros = (psihs > nonzero_th * psihs.max(axis=1)[:, None])
        avg_nonzeros = psihs_nonzeros.sum() / span
        if avg_nonzeros > nonzero_tol:
            continue
        psihs_peaks = np.where(psihs == psihs.max(axis=1)[:, None])
        joint_peak = np.argmax(np.prod(psihs, 0))
        if check_group(psihs_peaks, joint_peak, method, tol):
            break
    return i if (i < n_groups - 1) else None


def find_max_scale_alt(wavelet, N, min_cutoff=.1, max_cutoff=.8):
    """Max scale placing two bins symmetrically about the peak
    (cwt_utils.py:630-695)."""
    if max_cutoff <= 0 or min_cutoff <= 0:
        raise ValueError("`max_cutoff` and `min_cutoff` must be positive")
    elif max_cutoff <= min_cutoff:
        raise ValueError("must have `max_cutoff > min_cutoff`")
    wavelet = Wavelet._init_if_not_isinstance(wavelet)
    w_peak, peak = find_maximum(wavelet.fn)
    w_cutoff, _ = find_first_occurrence(wavelet.fn, value=min_cutoff * peak,
                                        step_start=0, step_limit=w_peak)
    w_ltp = np.arange(w_cutoff, w_peak, step=1 / N)
    div_size = (w_peak - w_ltp[:-1]) * 2
    n_divs = w_ltp[:-1] / div_size
    try:
        idx = np.where(np.diff(n_divs % 1) < -.8)[0][0]
    except IndexError:
        raise Exception("Failed to find suffciently-integer xi divisions; try "
                        "widening (min_cutoff, max_cutoff)")
    div_scale = div_size[idx + 1]
    w_1div = pi / (N / 2)
    return div_scale / w_1div
