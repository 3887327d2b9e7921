"""Inverses on the transform outputs: ``issq_cwt`` (_ssq_cwt.py:313-418),
``issq_stft`` (_ssq_stft.py:139-198) and the one-integral ``icwt``
(_cwt.py:323-494), with the admissibility constants ``adm_ssq`` /
``adm_cwt`` and their numeric integral ``integrate_analytic``
(utils/cwt_utils.py:28-63, 583-627).

The constants are host float64 arithmetic restated from the reference; the
column sums over Tx / Wx run in libssqk (``ssqk_real_rowsum``,
``ssqk_real_bandsum``), in NumPy's axis-0 order, so the float32 results
equal the reference's CPU results bit for bit.
"""
import numpy as np
from scipy import integrate

from . import _lib
from .wavelets import Wavelet

__all__ = ['issq_cwt', 'issq_stft', 'icwt', 'adm_ssq', 'adm_cwt', 'integrate_analytic']


def _min_neglect_idx(arr, th=1e-12):
    """algos.py:617-622: first index below th (else the last index)."""
    i = 0
    for i, x in enumerate(arr):
        if x < th:
            return i
    return i


def integrate_analytic(int_fn, nowarn=False):
    """Trapezoid integral over (0, inf) of a function that is zero for
    negative inputs, unimodal and decaying (utils/cwt_utils.py:583-627):
    1e-15..0.1 in log space, then 0.1..mxlim with the first mxlim in
    (1, 20, 80, 160) that leaves enough decay past the peak."""
    def _est_arr(mxlim, N):
        t = np.linspace(mxlim, .1, N, endpoint=False)[::-1].copy()
        arr = int_fn(t)
        max_idx = np.argmax(arr)
        min_neglect_idx = _min_neglect_idx(np.abs(arr[max_idx:]), th=1e-15) + max_idx
        return arr, t, min_neglect_idx

    t0 = np.logspace(-15, -1, 1000)
    int_nz = integrate.trapezoid(int_fn(t0), t0)
    for m, mxlim in zip([1, 1, 4, 8], [1, 20, 80, 160]):
        arr, t, min_neglect_idx = _est_arr(mxlim, N=10000 * m)
        if len(t) - min_neglect_idx > 1000 * m and np.sum(np.abs(arr)) > 1e-5:
            break
    else:
        if int_nz < 1e-5:
            raise Exception("Could not find converging or non-negligibly"
                            "-valued bounds of integration for `int_fn`")
        elif not nowarn:
            import warnings
            warnings.warn("Integrated only from 1e-15 to 0.1 in logspace")
    return integrate.trapezoid(arr[:min_neglect_idx], t[:min_neglect_idx]) + int_nz


def adm_ssq(wavelet):
    """Synchrosqueezing admissibility constant integral(conj(psih(w)) / w,
    w = 0..inf) (utils/cwt_utils.py:28-47)."""
    fn = Wavelet._init_if_not_isinstance(wavelet).fn
    Css = integrate_analytic(lambda w: np.conj(np.asarray(fn(w))) / w)
    return Css.real if abs(np.imag(Css)) < 1e-15 else Css


def adm_cwt(wavelet):
    """CWT admissibility constant integral(|psih(w)|^2 / w, w = 0..inf)
    (utils/cwt_utils.py:50-63)."""
    fn = Wavelet._init_if_not_isinstance(wavelet).fn

    def int_fn(w):
        p = np.asarray(fn(w))
        return np.conj(p) * p / w
    Cpsi = integrate_analytic(int_fn)
    return Cpsi.real if abs(np.imag(Cpsi)) < 1e-15 else Cpsi


def _device_tensor(T, what):
    import torch
    if not torch.is_tensor(T):
        T = torch.as_tensor(np.asarray(T))
    if not T.is_complex():
        raise TypeError("`%s` must be complex (got %s)" % (what, T.dtype))
    if T.dtype != torch.complex64:
        raise NotImplementedError("`%s`: only complex64 is supported on the GPU" % what)
    dev = T.device if T.is_cuda else torch.device('cuda', torch.cuda.current_device())
    return T.to(dev).contiguous()


def _component_args(cc, cw):
    """_process_component_inversion_args (_ssq_cwt.py:406-417)."""
    if cc is None and cw is None:
        return None, None, True
    cc, cw = np.asarray(cc), np.asarray(cw)
    if cc.ndim == 1:
        cc = cc.reshape(-1, 1)
    if cw.ndim == 1:
        cw = cw.reshape(-1, 1)
    return cc.astype('int32'), cw.astype('int32'), False


def _invert(T, cc, cw, scale, astensor):
    """Full (float32 [N], or [batch, N]) or curve-band (float64 [K + 1, N])
    inversion of T's real part, times scale."""
    import torch
    cc, cw, full = _component_args(cc, cw)
    if full:
        if T.ndim not in (2, 3):
            raise ValueError("`Tx` must be 2D (or 3D, a batch): got shape %s"
                             % (tuple(T.shape),))
        x = torch.empty(T.shape[:-2] + T.shape[-1:], dtype=torch.float32, device=T.device)
        _lib.real_rowsum(T, 0, T.shape[-2], scale, x)
    else:
        if T.ndim != 2:
            raise ValueError("curve-band inversion takes a 2D `Tx`")
        N = T.shape[-1]
        if cc.shape != cw.shape or cc.shape[0] != N:
            raise ValueError("`cc` and `cw` must both be [%d, K] (got %s, %s)"
                             % (N, cc.shape, cw.shape))
        ccd = torch.as_tensor(np.ascontiguousarray(cc), device=T.device)
        cwd = torch.as_tensor(np.ascontiguousarray(cw), device=T.device)
        x = torch.empty((cc.shape[1] + 1, N), dtype=torch.float64, device=T.device)
        _lib.real_bandsum(T, ccd, cwd, scale, x)
    return x if astensor else x.cpu().numpy()


def issq_cwt(Tx, wavelet='gmw', cc=None, cw=None, astensor=True):
    """Inverse synchrosqueezed CWT (_ssq_cwt.py:313-378): 2 / Css times the
    sum of Re Tx over frequency rows (cc = cw = None), or over curve bands
    cc +- cw per column (components, then the remainder)."""
    T = _device_tensor(Tx, 'Tx')
    wavelet = Wavelet._init_if_not_isinstance(wavelet)
    return _invert(T, cc, cw, 2 / adm_ssq(wavelet), astensor)


def issq_stft(Tx, window=None, cc=None, cw=None, n_fft=None, win_len=None, hop_len=1,
              modulated=True, astensor=True):
    """Inverse synchrosqueezed STFT (_ssq_stft.py:139-198): 2 / window[n//2]
    times the sum of Re Tx over frequency rows (or curve bands)."""
    from ._stft import get_window, _check_NOLA
    if not modulated:
        raise ValueError("inversion with `modulated == False` is unsupported.")
    if hop_len != 1:
        raise ValueError("inversion with `hop_len != 1` is unsupported.")
    T = _device_tensor(Tx, 'Tx')
    n_fft = n_fft or (T.shape[-2] - 1) * 2
    win_len = win_len or n_fft
    window = get_window(window, win_len, n_fft=n_fft)
    _check_NOLA(window, hop_len)
    if abs(np.argmax(window) - len(window) // 2) > 1:
        import warnings
        warnings.warn("`window` maximum not centered; results may be inaccurate.")
    return _invert(T, cc, cw, 2 / window[len(window) // 2], astensor)


def icwt(Wx, wavelet='gmw', scales='log-piecewise', nv=None, one_int=True, x_len=None,
         x_mean=0, padtype='reflect', rpadded=False, l1_norm=True, astensor=True):
    """One-integral inverse CWT (_cwt.py:323-494): sum over scales of
    Re Wx / norm(scale) times (2 / Css) ln(2^(1/nv)) (log scales) or
    (2 / Css) pi / 4 (linear), plus x_mean; log-piecewise scales invert each
    log segment and add them.  Wx: [na, N] or [batch, na, N]."""
    import torch
    from ._scales import process_scales, logscale_transition_idx
    W = _device_tensor(Wx, 'Wx')
    if not one_int and W.ndim == 3:
        raise NotImplementedError("batched `Wx` requires `one_int=True`.")
    if not one_int and rpadded:
        raise NotImplementedError("icwt(one_int=False, rpadded=True) is not supported")
    na, n = W.shape[-2], W.shape[-1]
    x_len = x_len or n
    if not (isinstance(scales, np.ndarray) or torch.is_tensor(scales)) and nv is None:
        nv = 32
    from ._cwt import process_gmw_wavelet
    wavelet = Wavelet._init_if_not_isinstance(process_gmw_wavelet(wavelet, l1_norm))
    scales, scaletype, _, nv = process_scales(scales, x_len, wavelet, nv=nv,
                                              get_params=True)
    scales = np.asarray(scales)
    assert len(scales) == na, "%s != %s" % (len(scales), na)
    if not one_int:
        return _icwt_2int(W, wavelet, scales, scaletype, nv, l1_norm, x_len, padtype, x_mean,
                          astensor)
    Css = adm_ssq(wavelet)
    x = torch.empty(W.shape[:-2] + W.shape[-1:], dtype=torch.float32, device=W.device)

    def segment(r0, r1, s, accumulate):
        # (the reference's recursion per log segment re-derives its scale
        # type and nv from the segment's scales)
        # norm(scale) of _icwt_norm (_cwt.py:485-494): L1 1 | scale, L2
        # scale**.5 | scale**1.5 (log | linear), in float32 as the reference
        _, stype, _, nv_s = process_scales(s, x_len, wavelet, get_params=True)
        s32 = np.asarray(s, dtype=np.float32).reshape(-1)
        if stype == 'log':
            norm = None if l1_norm else s32**.5
            k = (2 / Css) * np.log(2 ** (1 / nv_s))
        else:
            norm = s32 if l1_norm else s32**1.5
            k = (2 / Css) * np.pi / 4
        div = None if norm is None else torch.as_tensor(norm, device=W.device)
        _lib.real_rowsum(W, r0, r1, k, x, div=div, add=x_mean, accumulate=accumulate)

    if scaletype == 'log-piecewise':
        idx = logscale_transition_idx(scales)
        segment(0, idx, scales[:idx], False)
        segment(idx, na, scales[idx:], True)
    else:
        segment(0, na, scales, False)
    return x if astensor else x.cpu().numpy()


def _icwt_2int(W, wavelet, scales, scaletype, nv, l1_norm, x_len, padtype, x_mean, astensor):
    """Double-integral icwt (_cwt.py:456-476 and its constants,
    _cwt.py:437-452): x = Re sum_i ifft(psih_i fft(pad(Wx_i))) f_i,
    unpadded, f_i = (2 / Cpsi) ln(2^(1/nv)) / norm(s_i) (log) or
    (2 / Cpsi) (pi / 4) / norm(s_i) (linear), Cpsi = adm_cwt; a
    log-piecewise grid is two log segments, each adding x_mean.  The
    reference's (-1)^n on psih and its ifftshift cancel for an even padded
    length.  Device work: ssqk_pad_complex (pad, times f_i) and
    ssqk_cwt_adjoint of a plan of the padded length with the whole Nyquist
    bin (the reference's wavelet(scale=, N=))."""
    import torch
    from ._cwt import _padding, _wavelet_kernel_params
    from ._scales import process_scales, logscale_transition_idx
    na, n = W.shape[-2], W.shape[-1]
    Cpsi = adm_cwt(wavelet)

    def factors(s):
        _, stype, _, nv_s = process_scales(s, x_len, wavelet, get_params=True)
        s32 = np.asarray(s, dtype=np.float32).reshape(-1).astype(np.float64)
        if stype == 'log':
            k = (2 / Cpsi) * np.log(2 ** (1 / nv_s))
            norm = np.ones_like(s32) if l1_norm else s32**.5
        else:
            k = (2 / Cpsi) * np.pi / 4
            norm = s32 if l1_norm else s32**1.5
        return k / norm

    if scaletype == 'log-piecewise':
        idx = logscale_transition_idx(scales)
        f = np.concatenate([factors(scales[:idx]), factors(scales[idx:])])
        n_seg = 2
    else:
        f = factors(scales)
        n_seg = 1
    L, n1, pad = _padding(n, padtype)
    Wp = torch.empty((na, L), dtype=torch.complex64, device=W.device)
    f_d = torch.as_tensor(f.astype(np.float32), device=W.device)
    _lib.pad_complex(W, L, n1, pad, Wp, row_factor=f_d)
    scales64 = np.asarray(scales, dtype=np.float64).reshape(-1)
    plan = _lib.cwt_plan(W.device, L, L, 0, 'zero',
                         wavelet.kind | _lib.WAVELET_FULL_NYQUIST,
                         _wavelet_kernel_params(wavelet), scales64, 1.0)
    xp = torch.empty(L, dtype=torch.float32, device=W.device)
    plan.adjoint(xp, gWx=Wp)
    x = xp[n1:n1 + x_len]
    if x_mean:
        x = x + float(x_mean) * n_seg
    return x if astensor else x.cpu().numpy()
