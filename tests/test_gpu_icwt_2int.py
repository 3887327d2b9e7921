"""Double-integral icwt on the GPU (icwt(one_int=False), _cwt.py:437-476):
ssqk_pad_complex + ssqk_cwt_adjoint of a full-Nyquist plan against the
oracle's float64 restatement on the same Wx, for log, log-piecewise and
linear scales, both norms and the pad modes."""
import numpy as np
import pytest

from oracle import ssq_oracle as O
from test_gpu_cwt import rel_l2

pytestmark = pytest.mark.gpu


def _factors(scales, wav, n, l1_norm):
    from ssqueezepy_amd._inverse import adm_cwt
    from ssqueezepy_amd._scales import process_scales, logscale_transition_idx
    scales = np.asarray(scales)
    _, stype, _, nv = process_scales(scales, n, wav, get_params=True)

    def seg(s):
        _, st, _, nv_s = process_scales(s, n, wav, get_params=True)
        s64 = np.asarray(s, np.float32).reshape(-1).astype(np.float64)
        if st == 'log':
            return (2 / adm_cwt(wav)) * np.log(2 ** (1 / nv_s)) / (1 if l1_norm else s64**.5)
        return (2 / adm_cwt(wav)) * np.pi / 4 / (s64 if l1_norm else s64**1.5)
    if stype == 'log-piecewise':
        i = logscale_transition_idx(scales)
        return np.concatenate([seg(scales[:i]) * np.ones(i),
                               seg(scales[i:]) * np.ones(len(scales) - i)]), 2
    return seg(scales) * np.ones(len(scales)), 1


@pytest.mark.parametrize('case', [
    ('gmw', 'log-piecewise', True, 'reflect'),
    ('gmw', 'log', True, 'zero'),
    ('gmw', 'linear', True, 'symmetric'),
    ('morlet', 'log', False, 'reflect'),
    ('morlet', 'linear', False, 'wrap'),
])
def test_icwt_2int_against_oracle(case):
    import ssqueezepy_amd as sq
    from test_inverse_host import SIGNALS, mad_rms
    kind, scales, l1_norm, padtype = case
    wav = sq.Wavelet(('morlet', {'mu': 6.}) if kind == 'morlet' else ('gmw', {'beta': 8.}))
    pf = O.psih_fn('morlet', mu=6.) if kind == 'morlet' else O.psih_fn('gmw', beta=8.)
    x, ts = SIGNALS['lchirp'](2048)
    print(f"icwt one_int=False {case}", flush=True)
    wspec = ('morlet', {'mu': 6.}) if kind == 'morlet' else ('gmw', {'beta': 8.})
    Wx, s = sq.cwt(x.astype(np.float32), wspec, scales=scales, l1_norm=l1_norm, nv=32, t=ts)
    xr = sq.icwt(Wx, wspec, scales=scales, nv=32, one_int=False, l1_norm=l1_norm,
                 padtype=padtype, x_mean=0.25).cpu().numpy()
    f, n_seg = _factors(s.cpu().numpy(), wav, len(x), l1_norm)
    xo = O.icwt_2int(Wx.cpu().numpy(), pf, s.cpu().numpy().astype(np.float64), f,
                     padtype=padtype) + 0.25 * n_seg
    err = rel_l2(xr, xo)
    print(f"  rel_l2 {err:.3e}  mad_rms {mad_rms(x, xr - 0.25 * n_seg):.4f}", flush=True)
    assert err < 1e-5
    if scales != 'linear':
        assert mad_rms(x, xr - 0.25 * n_seg) < 0.1
