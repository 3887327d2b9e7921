"""Inverses on the GPU (ssqk_real_rowsum, ssqk_real_bandsum): bit for bit
against the oracle's NumPy restatement of the reference on the same Tx / Wx,
and the reference's reconstruction thresholds (tests/reconstruction_test.py)
on the float32 transforms."""
import numpy as np
import pytest

from oracle import ssq_oracle as O
from test_inverse_host import SIGNALS, mad_rms, _cases

pytestmark = pytest.mark.gpu

WAV8 = ('gmw', {'beta': 8.})


@pytest.mark.parametrize('name,scales', list(_cases()))
def test_issq_cwt_reconstructs(name, scales):
    """reconstruction_test.py:65-87 on the float32 GPU path: MAD/RMS < 0.1;
    and the inversion equals the oracle's NumPy expression on the same Tx
    bit for bit."""
    import torch
    import ssqueezepy_amd as sq
    x, ts = SIGNALS[name](2048)
    print(f"issq_cwt {name} {scales}", flush=True)
    Tx = sq.ssq_cwt(x.astype(np.float32), WAV8, scales=scales, nv=32, t=ts)[0]
    xrec = sq.issq_cwt(Tx, WAV8)
    assert xrec.dtype == torch.float32 and xrec.shape == (len(x),)
    xr = xrec.cpu().numpy()
    assert mad_rms(x, xr) < 0.1, mad_rms(x, xr)
    ref = O.issq_full(Tx.cpu().numpy(), 2 / sq.adm_ssq(WAV8))
    assert np.array_equal(xr, ref)


def test_issq_and_icwt_log_piecewise():
    """reconstruction_test.py:113-123: echirp(1024), default GMW,
    log-piecewise: issq_cwt and icwt MAD/RMS < 0.02."""
    import ssqueezepy_amd as sq
    x, ts = SIGNALS['echirp'](1024)
    print("issq/icwt log-piecewise", flush=True)
    Tx, Wx, _, scales = sq.ssq_cwt(x.astype(np.float32), 'gmw', scales='log-piecewise', t=ts)
    e1 = mad_rms(x, sq.issq_cwt(Tx, 'gmw').cpu().numpy())
    e2 = mad_rms(x, sq.icwt(Wx, 'gmw', scales=scales.cpu().numpy()).cpu().numpy())
    assert e1 < .02 and e2 < .02, (e1, e2)


def test_icwt_reconstructs_and_matches_oracle():
    """reconstruction_test.py:90-110 (l1_norm=True) on the GPU: MAD/RMS of
    icwt(cwt(x)) < 0.1; the log-piecewise inversion (two segments, scale
    factors per segment) equals the oracle's NumPy expression bit for bit."""
    import ssqueezepy_amd as sq
    from ssqueezepy_amd._scales import process_scales, logscale_transition_idx
    for name in SIGNALS:
        x, ts = SIGNALS[name](2048)
        stype = 'log:maximal' if name in ('low_freqs', 'high_freqs') else 'log'
        print(f"icwt {name}", flush=True)
        Wx, _ = sq.cwt(x.astype(np.float32), WAV8, scales=stype, nv=32, t=ts)
        xr = sq.icwt(Wx, WAV8, scales=stype, nv=32).cpu().numpy()
        assert mad_rms(x, xr) < 0.1, (name, mad_rms(x, xr))
    x, ts = SIGNALS['lchirp'](3000)
    Wx, scales = sq.cwt(x.astype(np.float32), 'gmw', t=ts)
    s = scales.cpu().numpy()
    xr = sq.icwt(Wx, 'gmw', scales=s, x_mean=0.25).cpu().numpy()
    Css = sq.adm_ssq('gmw')
    idx = logscale_transition_idx(s)
    W = Wx.cpu().numpy()
    parts = []
    for r0, r1 in ((0, idx), (idx, len(s))):
        nv = process_scales(s[r0:r1], 3000, sq.Wavelet('gmw'), get_params=True)[3]
        p = W[r0:r1].real.sum(axis=0)
        p *= (2 / Css) * np.log(2 ** (1 / nv))
        p += 0.25
        parts.append(p)
    assert np.array_equal(xr, parts[0] + parts[1])


def test_batched_inverse_equals_per_signal():
    import torch
    import ssqueezepy_amd as sq
    xs = np.stack([SIGNALS['lchirp'](4000)[0], SIGNALS['echirp'](4000)[0]]).astype(np.float32)
    print("batched issq_cwt", flush=True)
    Tx = sq.ssq_cwt(xs, 'gmw')[0]
    xb = sq.issq_cwt(Tx, 'gmw')
    assert xb.shape == (2, 4000)
    for b in range(2):
        assert torch.equal(xb[b], sq.issq_cwt(Tx[b].contiguous(), 'gmw'))


def test_component_inversion_matches_oracle():
    """Curve bands around the ridge of each column (one band per tone of a
    two-tone signal, -1 where a curve is absent, bands clipped at the first
    and last rows): components and remainder equal the oracle's restatement
    of _invert_components bit for bit, and add up to the full inversion."""
    import ssqueezepy_amd as sq
    N = 3000
    n = np.arange(N)
    x = (np.cos(2 * np.pi * 0.05 * n) + 0.5 * np.cos(2 * np.pi * 0.2 * n)).astype(np.float32)
    print("issq_cwt components", flush=True)
    Tx = sq.ssq_cwt(x, 'gmw')[0]
    T = Tx.cpu().numpy()
    rows = T.shape[0]
    a = np.abs(T)
    half = rows // 2
    c0 = a[:half].argmax(axis=0)
    c1 = half + a[half:].argmax(axis=0)
    cc = np.stack([c0, c1], axis=1).astype(np.int32)
    cw = np.full_like(cc, 3)
    cc[::11, 1] = -1
    cc[:5, 0] = 0                     # clipped at row 0
    cc[5:9, 1] = rows - 1             # clipped at the last row
    Css = sq.adm_ssq('gmw')
    xg = sq.issq_cwt(Tx, 'gmw', cc, cw).cpu().numpy()
    xo = O.invert_components(T, cc, cw, 2 / Css)
    assert xg.dtype == np.float64 and xg.shape == (3, N)
    assert np.array_equal(xg, xo)
    full = sq.issq_cwt(Tx, 'gmw').cpu().numpy().astype(np.float64)
    assert np.allclose(xg.sum(axis=0), full, rtol=0, atol=1e-5 * np.abs(full).max())


@pytest.mark.parametrize('N', [128, 129])
@pytest.mark.parametrize('n_fft', [64, 128])
@pytest.mark.parametrize('window_scaling', [1., .5])
def test_issq_stft_reconstructs(N, n_fft, window_scaling):
    """reconstruction_test.py:185-206 (n_fft 120 / 121 there: the GPU path
    takes powers of two): MAE of issq_stft(ssq_stft(x)) < 0.1, and the
    inversion equals the oracle's NumPy expression bit for bit."""
    import ssqueezepy_amd as sq
    from ssqueezepy_amd._stft import get_window
    x = np.random.default_rng(N + n_fft).standard_normal(N).astype(np.float32)
    window = None
    if window_scaling != 1:
        window = get_window(None, win_len=n_fft, n_fft=n_fft) * window_scaling
    print(f"issq_stft N={N} n_fft={n_fft} scaling={window_scaling}", flush=True)
    Sx = sq.ssq_stft(x, window=window, n_fft=n_fft)[0]
    xr = sq.issq_stft(Sx, window=window, n_fft=n_fft).cpu().numpy()
    assert len(xr) == N
    assert np.abs(x - xr).mean() < 0.1
    w = get_window(window, n_fft, n_fft=n_fft)
    assert np.array_equal(xr, O.issq_full(Sx.cpu().numpy(), 2 / w[len(w) // 2]))
