"""The oracle is pinned before anything is compared with it: against the
golden vectors written from the reference (tests/golden/make_golden.py) and
against what the mathematics fixes (brute-force DFTs, closed forms)."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import ssq_oracle as O

EPS32 = np.finfo(np.float32).eps


def _grid_of(e2e, key):
    scales = e2e[key + '/scales'].astype(np.float64)
    freqs = e2e[key + '/ssq_freqs'][::-1].copy()     # fixtures hold the flipped grid
    return scales, freqs


def _const_log(scales):
    nv = 1 / np.diff(np.log2(scales))
    nv = np.concatenate([nv[:1], nv])
    return np.log(2) / nv


# ------------------------------------------------------------ math pins
def test_cwt_matches_bruteforce_dft_convolution():
    rng = np.random.default_rng(5)
    x = rng.standard_normal(24)
    scales = np.array([1.1, 2.3, 5.0])
    psih = O.psih_fn('gmw', gamma=3., beta=60.)
    Wx, dWx = O.cwt(x, psih, scales, dt=0.5)
    Wr, dWr = O.cwt(x, psih, scales, dt=0.5, rpadded=True)
    xp, L, n1, _ = O.pad_signal(x, 'reflect')
    n = np.arange(L)
    xi = O.xi_grid(L)
    for i, s in enumerate(scales):
        H = psih(s * xi)
        H[L // 2] /= 2
        # y[n] = (1/L) sum_k H[k] X[k] e^{2 pi i k n / L} with X the DFT of xp
        X = np.array([np.sum(xp * np.exp(-2j * np.pi * k * n / L)) for k in range(L)])
        E = np.exp(2j * np.pi * np.outer(n, n) / L)
        y = E @ (H * X) / L
        yd = E @ (H * X * 1j * xi / 0.5) / L
        assert np.allclose(Wx[i], y[n1:n1 + len(x)], atol=1e-12)
        assert np.allclose(dWx[i], yd[n1:n1 + len(x)], atol=1e-11)
        # rpadded=True: every column of the circular convolution
        assert np.allclose(Wr[i], y, atol=1e-12) and np.allclose(dWr[i], yd, atol=1e-11)


def test_wavelet_closed_forms():
    # GMW bandpass: peak value 2 at wc = (beta/gamma)^(1/gamma) (_gmw.py:35-37)
    for g, b in [(3., 60.), (3., 12.), (2., 5.)]:
        wc = (b / g)**(1 / g)
        w = np.linspace(wc * 0.5, wc * 1.5, 20001)
        v = O.gmw_psih(w, g, b)
        assert abs(v.max() - 2) < 1e-7 and abs(w[np.argmax(v)] - wc) < 1e-3
        assert O.gmw_psih(np.array([-1., 0.]), g, b).tolist() == [0., 0.]
    # Morlet: psih(0) = 0 exactly (admissibility term), peak near mu
    assert abs(O.morlet_psih(np.array([0.]), 6.)[0]) < 1e-15
    w = np.linspace(3, 9, 60001)
    assert abs(w[np.argmax(O.morlet_psih(w, 6.))] - 6) < 1e-3


def test_gmw_time_domain_l1_norm_is_two():
    # tests/gmw_test.py:59-81: the bandpass GMW's time-domain wavelet
    # ifft(psih * (-1)^n) has L1 norm 2 within 1e-3
    for g, b in [(3., 60.), (4., 80.)]:
        for scale in (2., 3.):
            N = 512
            psih = O.gmw_psih(scale * O.xi_grid(N), g, b)
            psi = np.fft.ifft(psih * (-1)**np.arange(N))
            assert abs(np.abs(psi).sum() - 2) < 1e-3


def test_phase_transform_closed_form():
    rng = np.random.default_rng(0)
    Wx = rng.standard_normal((7, 50)) * (1 + 2j)
    dWx = rng.standard_normal((7, 50)) * (2 - 1j)
    w = O.phase_transform(Wx, dWx, 1e-2)
    ref = np.abs((dWx / Wx).imag / (2 * np.pi))
    ref[np.abs(Wx) < 1e-2] = np.inf
    assert np.allclose(w, ref, rtol=1e-12)


def test_squeeze_bruteforce_and_order():
    rng = np.random.default_rng(1)
    Wx = rng.standard_normal((9, 13)) + 1j * rng.standard_normal((9, 13))
    k = rng.integers(-1, 9, size=(9, 13))
    const = rng.random(9)
    Tx = O.ssqueeze(Wx, k, const)
    ref = np.zeros_like(Tx)
    for i in range(9):
        for j in range(13):
            if k[i, j] >= 0:
                ref[k[i, j], j] += Wx[i, j] * const[i]
    assert np.array_equal(Tx, ref)


def test_bins_round_half_even_and_clamps():
    p = dict(kind='lin', vmin=0., dv=1.)
    w = np.array([[0.5, 1.5, 2.5, -3., 99., np.inf]])
    k = O.squeeze_bins(w, p, omax=5, flipud=False)
    assert k.tolist() == [[0, 2, 2, 0, 5, -1]]
    assert O.squeeze_bins(w, p, omax=5, flipud=True).tolist() == [[5, 3, 3, 5, 0, -1]]


def test_stft_matches_dft_definition():
    rng = np.random.default_rng(3)
    x = rng.standard_normal(40)
    n = 16
    win = np.hanning(n + 1)[:-1]
    dwin = rng.standard_normal(n)
    Sx, dSx = O.stft(x, win, dwin, n, hop=3, fs=2.)
    xp, _, _, _ = O.pad_signal(x, 'reflect', padlength=len(x) + n - 1)
    for f in range(Sx.shape[1]):
        fr = xp[3 * f:3 * f + n]
        shift = np.roll(fr * win, -(n // 2))     # ifftshift of the frame*window
        assert np.allclose(Sx[:, f], np.fft.rfft(shift), atol=1e-12)
        assert np.allclose(dSx[:, f], np.fft.rfft(np.roll(fr * dwin * 2., -(n // 2))),
                           atol=1e-12)


# ---------------------------------------------------------- golden pins
@pytest.mark.parametrize('key,kind,params,logscale,fused', [
    ('gmw64_lp', 'gmw', {}, True, True),
    ('morlet64_lp_log', 'morlet', {'mu': 6.}, True, True),
    ('gmw64_log', 'gmw', {}, True, True),
    ('gmw64_lp_noflip', 'gmw', {}, True, True),
])
def test_oracle_ssq_cwt_vs_reference_golden(golden, key, kind, params, logscale, fused):
    e = golden('e2e_cwt')
    x = e['x'].astype(np.float64)
    scales, freqs = _grid_of(e, key)
    flip = not key.endswith('noflip')
    if flip is False:
        freqs = e[key + '/ssq_freqs'][::-1].copy()
    const = _const_log(scales)
    Tx, Wx, dWx, w, k = O.ssq_cwt(x, O.psih_fn(kind, **params), scales, freqs, const,
                                  logscale, 10 * np.finfo(np.float64).eps,
                                  flipud=flip)
    tol = 1e-10 if e[key + '/Wx'].dtype == np.complex128 else 2e-6
    assert rel_l2(Wx, e[key + '/Wx']) < tol
    assert rel_l2(Tx, e[key + '/Tx']) < max(tol, 1e-9)
    if key + '/dWx' in e:
        assert rel_l2(dWx, e[key + '/dWx']) < 1e-10
        wg = e[key + '/w']
        fin = np.isfinite(wg)
        assert np.array_equal(fin, np.isfinite(w))
        assert np.allclose(w[fin], wg[fin], rtol=1e-8, atol=0)
    assert float(e[key + '/Tx_fused_maxdiff']) < 1e-12


def test_oracle_linear_scales_vs_golden(golden):
    e = golden('e2e_cwt')
    key = 'gmw64_lin'
    x = e['x'].astype(np.float64)
    scales, freqs = _grid_of(e, key)
    const = (scales[1] - scales[0]) / scales
    Tx, Wx, *_ = O.ssq_cwt(x, O.psih_fn('gmw'), scales, freqs, const, False,
                           10 * np.finfo(np.float64).eps)
    assert rel_l2(Wx, e[key + '/Wx']) < 2e-6
    assert rel_l2(Tx, e[key + '/Tx']) < 2e-6


def test_oracle_batched_vs_golden(golden):
    e = golden('e2e_cwt')
    scales, freqs = _grid_of(e, 'batch')
    Tx, Wx, *_ = O.ssq_cwt(e['xb'].astype(np.float64), O.psih_fn('gmw'), scales, freqs,
                           _const_log(scales), True, 10 * np.finfo(np.float64).eps)
    assert rel_l2(Wx, e['batch/Wx']) < 2e-6
    assert rel_l2(Tx, e['batch/Tx']) < 2e-6


@pytest.mark.parametrize('key,window,n_fft,fs,flip', [
    ('hann64_64', 'hann', 64, 1., False),
    ('hann32_fs', 'hann', 32, 2., True),
])
def test_oracle_ssq_stft_vs_reference_golden(golden, key, window, n_fft, fs, flip):
    import scipy.signal as sig
    e = golden('e2e_stft')
    x = e['x'].astype(np.float64)
    win = sig.get_window(window, n_fft, fftbins=True)
    xi = O.xi_grid(n_fft)
    xi[n_fft // 2] = 0
    dwin = np.fft.ifft(np.fft.fft(win) * 1j * xi).real
    Tx, Sx, dSx, w, k, Sfs = O.ssq_stft(x, win, dwin, n_fft,
                                        10 * np.finfo(np.float64).eps, fs=fs,
                                        flipud=flip)
    assert rel_l2(Sx, e[key + '/Sx']) < 1e-12
    assert rel_l2(dSx, e[key + '/dSx']) < 1e-12
    assert rel_l2(Tx, e[key + '/Tx']) < 1e-12
    assert np.allclose(Sfs, e[key + '/Sfs'])
    wg = e[key + '/w']
    fin = np.isfinite(wg)
    assert np.allclose(w[fin], wg[fin], rtol=1e-9)


def test_oracle_stft_hop3_vs_golden(golden):
    import scipy.signal as sig
    e = golden('e2e_stft')
    win = sig.get_window('hann', 64, fftbins=True)
    xi = O.xi_grid(64)
    xi[32] = 0
    dwin = np.fft.ifft(np.fft.fft(win) * 1j * xi).real
    Sx, dSx = O.stft(e['x'].astype(np.float64), win, dwin, 64, hop=3)
    assert rel_l2(Sx, e['hop3/Sx']) < 1e-12
    assert rel_l2(dSx, e['hop3/dSx']) < 1e-12


@pytest.mark.parametrize('stype', ['log-piecewise', 'log', 'linear'])
@pytest.mark.parametrize('flip', [0, 1])
def test_oracle_squeeze_vs_reference_golden(golden, stype, flip):
    s = golden('squeeze')
    Wx, dWx, w, const = s['Wx'], s['dWx'], s['w'], s['const']
    f = s[stype + '/ssq_freqs']
    log = stype.startswith('log')
    na = len(Wx)
    k = O.squeeze_bins(w, O.bin_params(f, log), na - 1, bool(flip))
    assert np.allclose(O.ssqueeze(Wx, k, const), s[f'{stype}/{flip}/onfly'], atol=1e-12)
    g = 1e-2
    wf = np.abs((dWx.imag * Wx.real - dWx.real * Wx.imag)
                / ((Wx.real**2 + Wx.imag**2) * O.TWO_PI))
    wf[~(np.abs(Wx) > g)] = np.inf
    k = O.squeeze_bins(wf, O.bin_params(f / na / 4, log), na - 1, bool(flip))
    assert np.allclose(O.ssqueeze(Wx, k, const), s[f'{stype}/{flip}/fast'], atol=1e-12)


def test_oracle_phase_and_stft_squeeze_vs_golden(golden):
    s = golden('squeeze')
    Wx, dWx, const, Sfs = s['Wx'], s['dWx'], s['const'], s['stft/Sfs']
    assert np.allclose(O.phase_transform(Wx, dWx, 1e-2), s['phase_cwt'])
    assert np.allclose(O.phase_transform(Wx, dWx, 1e-2, Sfs=Sfs), s['phase_stft'])
    A, B, C, D = dWx.real, dWx.imag, Wx.real, Wx.imag
    wf = np.abs(Sfs[:, None] - (B * C - A * D) / ((C**2 + D**2) * O.TWO_PI))
    wf[~(np.abs(Wx) > 1e-2)] = np.inf
    k = O.squeeze_bins(wf, O.bin_params(Sfs, False), len(Wx) - 1, False)
    assert np.allclose(O.ssqueeze(Wx, k, const), s['stft/fast'], atol=1e-12)


def test_oracle_c1_samples_vs_golden(golden):
    c = golden('c1')
    scales = c['scales'].astype(np.float64)
    freqs = c['ssq_freqs'][::-1].copy()
    Tx, Wx, dWx, w, k = O.ssq_cwt(c['x'].astype(np.float64), O.psih_fn('morlet', mu=6.),
                                  scales, freqs, _const_log(scales), True,
                                  10 * np.finfo(np.float64).eps, fused=False)
    cols = c['cols']
    assert rel_l2(Wx[:, cols], c['Wx_cols']) < 1e-12
    assert rel_l2(np.linalg.norm(Tx, axis=1), c['Tx_rownorm']) < 1e-12
    assert rel_l2(Tx[:, cols], c['Tx_cols']) < 1e-12
    fin = np.isfinite(c['w_cols'])
    assert np.allclose(w[:, cols][fin], c['w_cols'][fin], rtol=1e-9)
