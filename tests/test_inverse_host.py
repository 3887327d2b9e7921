"""Inverses on the host side (no GPU): the admissibility constants against
closed forms and an independent integrator, and the oracle's inversions
against the reference's reconstruction thresholds
(tests/reconstruction_test.py:21-123) on the float64 chain."""
import numpy as np
import pytest
from scipy import integrate

from oracle import ssq_oracle as O


def _t(lo, hi, N):
    return np.linspace(lo, hi, N, endpoint=False)


def _cos_f(freqs, N=128):
    return np.concatenate([np.cos(2 * np.pi * f * _t(i, i + 1, N))
                           for i, f in enumerate(freqs)])


def _freqs(N, freqs):
    x = _cos_f(freqs, N // len(freqs))
    return x, _t(0, len(x) / N, len(x))


# the reference's reconstruction test signals (reconstruction_test.py:35-56)
SIGNALS = {
    'echirp': lambda N: (np.cos(2 * np.pi * 3 * np.exp(_t(0, 10, N) / 3)), _t(0, 10, N)),
    'lchirp': lambda N: (np.cos(np.pi * _t(0, 10, N)**2), _t(0, 10, N)),
    'fast_transitions': lambda N: _freqs(N, np.array([N / 100, N / 200, N / 3, N / 20,
                                                      N / 3 - 1, N / 50, N / 4,
                                                      N / 150]) / 8),
    'low_freqs': lambda N: _freqs(N, [.3, .3, 1, 1, 2, 2]),
    'high_freqs': lambda N: _freqs(N, np.array([N / 2, N / 2 - 1, N / 4, N / 3]) / 4),
}


def mad_rms(x, xrec):
    return np.mean(np.abs(x - xrec)) / np.sqrt(np.mean(x**2))


def test_adm_ssq_and_adm_cwt():
    """integrate_analytic restated: within its trapezoid error of the GMW
    closed forms, and of scipy's adaptive quadrature for Morlet."""
    import ssqueezepy_amd as sq
    from ssqueezepy_amd._inverse import adm_ssq, adm_cwt
    for beta in (60., 8.):
        w = sq.Wavelet(('gmw', {'beta': beta}))
        assert abs(adm_ssq(w) / O.gmw_adm_ssq(3., beta) - 1) < 1e-5
        assert abs(adm_cwt(w) / O.gmw_adm_cwt(3., beta) - 1) < 1e-5
    fn = sq.Wavelet(('morlet', {'mu': 6.})).fn
    q = integrate.quad(lambda v: fn(np.array([v]))[0] / v, 0, 40, limit=400)[0]
    assert abs(adm_ssq(('morlet', {'mu': 6.})) / q - 1) < 1e-6


def _cases():
    for name in SIGNALS:
        for scales in ('log', 'log-piecewise', 'linear'):
            if name == 'low_freqs':
                if scales == 'linear':
                    continue
                scales = scales + ':maximal'
            yield name, scales


@pytest.mark.parametrize('name,scales', list(_cases()))
def test_oracle_issq_cwt_reconstructs(name, scales):
    """reconstruction_test.py:65-87: MAD/RMS of issq_cwt(ssq_cwt(x)) < 0.1
    (GMW beta=8, float64)."""
    import ssqueezepy_amd as sq
    from ssqueezepy_amd._ssq_cwt import ssq_cwt_grid
    from ssqueezepy_amd._inverse import adm_ssq
    x, ts = SIGNALS[name](2048)
    wav = sq.Wavelet(('gmw', {'beta': 8., 'dtype': 'float64'}))
    g = ssq_cwt_grid(len(x), wav, scales=scales, nv=32, t=ts)
    Tx = O.ssq_cwt(x, O.psih_fn('gmw', beta=8.), g['scales64'], g['ssq_freqs'],
                   g['const'], g['bp'].type != 0, 10 * np.finfo(np.float64).eps,
                   dt=g['dt'])[0]
    xrec = O.issq_full(Tx, 2 / adm_ssq(wav))
    assert mad_rms(x, xrec) < 0.1, (name, scales, mad_rms(x, xrec))


def test_oracle_components_add_up_to_the_full_inversion():
    """Curve bands that do not overlap: components + remainder = the full
    inversion (float64 against the float32-sum remainder's rounding)."""
    rng = np.random.default_rng(0)
    rows, N = 60, 300
    Tx = (rng.standard_normal((rows, N)) + 1j * rng.standard_normal((rows, N))
          ).astype(np.complex64)
    cc = np.stack([rng.integers(5, 20, N), rng.integers(35, 55, N)], axis=1).astype(np.int32)
    cw = np.stack([rng.integers(0, 5, N), rng.integers(0, 5, N)], axis=1).astype(np.int32)
    cc[::7, 0] = -1
    x = O.invert_components(Tx, cc, cw, 1.0)
    full = Tx.astype(np.complex128).real.sum(axis=0)
    assert np.allclose(x.sum(axis=0), full, rtol=0, atol=1e-5)
    # a band of every row is the full inversion, nothing left over
    allb = O.invert_components(Tx, np.full((N, 1), rows // 2, np.int32),
                               np.full((N, 1), rows, np.int32), 1.0)
    assert np.array_equal(allb[0], full) and not allb[1].any()


def test_oracle_icwt_reconstructs():
    """reconstruction_test.py:90-110 (l1_norm=True): MAD/RMS of
    icwt(cwt(x)) < 0.1, one-integral inverse, log scales."""
    import ssqueezepy_amd as sq
    from ssqueezepy_amd._inverse import adm_ssq
    from ssqueezepy_amd._scales import process_scales
    wav = sq.Wavelet(('gmw', {'beta': 8., 'dtype': 'float64'}))
    for name in SIGNALS:
        x, ts = SIGNALS[name](2048)
        stype = 'log:maximal' if name in ('low_freqs', 'high_freqs') else 'log'
        scales, _, _, nv = process_scales(stype, len(x), wav, nv=32, get_params=True)
        s64 = np.asarray(scales, dtype=np.float64).reshape(-1)
        Wx, _ = O.cwt(x, O.psih_fn('gmw', beta=8.), s64, dt=ts[1] - ts[0])
        k = (2 / adm_ssq(wav)) * np.log(2 ** (1 / nv))
        xrec = O.icwt_1int(Wx, [(0, len(s64), k, None)])
        assert mad_rms(x, xrec) < 0.1, (name, mad_rms(x, xrec))


def test_adm_coefficients_stay_away_from_zero():
    """adm_coef_test.py:17-36: the numeric integrals of 200 Morlet
    wavelets (mu = 4 .. 30) and of GMW beta = 2 .. 120 stay above 1e-3
    (an unstable integration gives values near zero)."""
    from ssqueezepy_amd._inverse import adm_ssq, adm_cwt
    for wav in ([('morlet', {'mu': mu}) for mu in np.linspace(4, 30, 200)] +
                [('gmw', {'beta': b}) for b in np.linspace(2, 120, 40)]):
        assert adm_cwt(wav) > 1e-3 and adm_ssq(wav) > 1e-3, wav


def test_oracle_l2_cwt_equal_energy_per_scale():
    """cwt(l1_norm=False) (_cwt.py:307-311): the L2-normalized wavelet of
    every scale has the same time-domain energy, so an impulse's transform
    has equal row energies wherever the scale's band and support fit the
    grid; the L1 rows fall as 1/scale."""
    import ssqueezepy_amd as sq
    from ssqueezepy_amd._scales import process_scales
    N = 4096
    x = np.zeros(N)
    x[N // 2] = 1
    wav = sq.Wavelet(('morlet', {'mu': 6., 'dtype': 'float64'}))
    s = np.asarray(process_scales('log', N, wav, nv=16), dtype=np.float64).ravel()
    mid = slice(len(s) // 4, 3 * len(s) // 4)
    e = {}
    for l1 in (True, False):
        Wx, _ = O.cwt(x, O.psih_fn('morlet', mu=6.), s, padtype='zero', derivative=False,
                      l1_norm=l1)
        e[l1] = (np.abs(Wx[mid])**2).sum(-1)
    assert e[False].max() / e[False].min() - 1 < 1e-9
    assert np.allclose(e[True] * s[mid], e[True][0] * s[mid][0], rtol=1e-9)


def test_oracle_icwt_l2_reconstructs():
    """reconstruction_test.py:88-107 with l1_norm=False (Morlet: the
    reference's L2 GMW is float64-only): MAD/RMS of icwt(cwt(x)) < 0.1 with
    norm(scale) = scale**.5 on log scales."""
    import ssqueezepy_amd as sq
    from ssqueezepy_amd._inverse import adm_ssq
    from ssqueezepy_amd._scales import process_scales
    wav = sq.Wavelet(('morlet', {'mu': 6., 'dtype': 'float64'}))
    for name in SIGNALS:
        x, ts = SIGNALS[name](2048)
        stype = 'log:maximal' if name in ('low_freqs', 'high_freqs') else 'log'
        scales, _, _, nv = process_scales(stype, len(x), wav, nv=32, get_params=True)
        s64 = np.asarray(scales, dtype=np.float64).reshape(-1)
        Wx, _ = O.cwt(x, O.psih_fn('morlet', mu=6.), s64, dt=ts[1] - ts[0], derivative=False,
                      l1_norm=False)
        k = (2 / adm_ssq(wav)) * np.log(2 ** (1 / nv))
        xrec = O.icwt_1int(Wx, [(0, len(s64), k, np.asarray(scales, np.float32).ravel()**.5)])
        assert mad_rms(x, xrec) < 0.1, (name, mad_rms(x, xrec))
        Wx1, _ = O.cwt(x, O.psih_fn('morlet', mu=6.), s64, dt=ts[1] - ts[0],
                       derivative=False)
        assert np.allclose(Wx, Wx1 * np.sqrt(s64)[:, None], rtol=1e-12, atol=1e-14)


def test_l1_norm_false_wavelet_rules():
    """_process_gmw_wavelet (_cwt.py:497-514): GMW takes the 'energy' norm
    (float64 in the reference; not on the float32 path), a GMW L1 Wavelet
    with l1_norm=False is a ValueError, Morlet passes through."""
    import ssqueezepy_amd as sq
    from ssqueezepy_amd._cwt import process_gmw_wavelet
    assert process_gmw_wavelet('gmw', False) == ('gmw', {'norm': 'energy'})
    assert process_gmw_wavelet(('gmw', {'beta': 8}), False) == (
        'gmw', {'beta': 8, 'norm': 'energy'})
    assert process_gmw_wavelet('morlet', False) == 'morlet'
    with pytest.raises(ValueError):
        process_gmw_wavelet(sq.Wavelet('gmw'), False)
    with pytest.raises(NotImplementedError):
        sq.cwt(np.zeros(64, np.float32), 'gmw', l1_norm=False)


def test_oracle_icwt_2int_reconstructs():
    """Double-integral icwt (_cwt.py:456-476): MAD/RMS of icwt(cwt(x)) <
    0.1 with Cpsi = adm_cwt for the chirps and the fast transitions (GMW
    beta = 8 and Morlet mu = 6, log scales); an error in the constant, the
    wavelet or the shifts fails it."""
    import ssqueezepy_amd as sq
    from ssqueezepy_amd._inverse import adm_cwt
    from ssqueezepy_amd._scales import process_scales
    for wav, pf in [(sq.Wavelet(('gmw', {'beta': 8., 'dtype': 'float64'})),
                     O.psih_fn('gmw', beta=8.)),
                    (sq.Wavelet(('morlet', {'mu': 6., 'dtype': 'float64'})),
                     O.psih_fn('morlet', mu=6.))]:
        for name in ('echirp', 'lchirp', 'fast_transitions'):
            x, ts = SIGNALS[name](2048)
            scales, _, _, nv = process_scales('log', len(x), wav, nv=32, get_params=True)
            s64 = np.asarray(scales, dtype=np.float64).reshape(-1)
            Wx, _ = O.cwt(x, pf, s64, dt=ts[1] - ts[0], derivative=False)
            k = (2 / adm_cwt(wav)) * np.log(2 ** (1 / nv))
            xrec = O.icwt_2int(Wx, pf, s64, np.full(len(s64), k))
            assert mad_rms(x, xrec) < 0.1, (name, mad_rms(x, xrec))
            assert mad_rms(x, 0.5 * xrec) > 0.3       # the constant matters
