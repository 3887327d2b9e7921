"""cwt(l1_norm=False) and icwt(l1_norm=False) on the GPU (_cwt.py:307-311,
485-494): the L2-normalized Morlet transform against the oracle's float64
one, its gradient, and the reference's reconstruction threshold
(tests/reconstruction_test.py:88-107) through both inverses' norms."""
import numpy as np
import pytest

from oracle import ssq_oracle as O
from test_gpu_cwt import chirp_tone_noise, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('N', [256, 3000, 40000])
def test_cwt_l2_against_oracle(N):
    """Every path of the plan (band, overlap-save, four-step) with the
    sqrt(scale) weights in its wavelet tables; Wx and dWx."""
    import ssqueezepy_amd as sq
    x = chirp_tone_noise(N, 7)
    print(f"cwt l1_norm=False N={N}", flush=True)
    wav = ('morlet', {'mu': 6.})
    Wx, scales, dWx = sq.cwt(x, wav, l1_norm=False, derivative=True)
    s64 = scales.cpu().numpy().astype(np.float64)
    Wo, dWo = O.cwt(x.astype(np.float64), O.psih_fn('morlet', mu=6.), s64, l1_norm=False)
    assert rel_l2(Wx.cpu().numpy(), Wo) < 1e-5
    assert rel_l2(dWx.cpu().numpy(), dWo) < 1e-5
    W1, _ = sq.cwt(x, wav)
    ratio = np.abs(Wx.cpu().numpy()).sum(-1) / np.abs(W1.cpu().numpy()).sum(-1)
    assert np.allclose(ratio, np.sqrt(s64), rtol=1e-4)


def test_cwt_l2_gradient_dot_product():
    """The gradient of Re <y, Wx(x)> is the L2 transform's adjoint A^H y:
    <grad, x> = Re <y, A x>."""
    import torch
    import ssqueezepy_amd as sq
    N = 3000
    gen = torch.Generator(device='cuda').manual_seed(2)
    x = torch.randn(N, device='cuda', generator=gen).requires_grad_(True)
    print("cwt l1_norm=False gradient", flush=True)
    Wx, _ = sq.cwt(x, 'morlet', l1_norm=False)
    y = torch.randn(Wx.shape, dtype=torch.complex64, device='cuda', generator=gen)
    loss = (Wx.real * y.real + Wx.imag * y.imag).double().sum()
    loss.backward()
    lhs = float((x.grad.double() * x.detach().double()).sum())
    L = float(loss.detach())
    assert abs(lhs - L) <= 1e-5 * abs(L) + 1e-6


def test_icwt_l2_reconstructs():
    """reconstruction_test.py:88-107 with l1_norm=False (Morlet, mu=6): MAD/RMS
    of icwt(cwt(x)) < 0.1, and icwt equals the oracle's sum with
    norm(scale) = scale**.5."""
    import ssqueezepy_amd as sq
    from ssqueezepy_amd._inverse import adm_ssq
    from ssqueezepy_amd._scales import process_scales
    from test_inverse_host import SIGNALS, mad_rms
    for name in SIGNALS:
        x, ts = SIGNALS[name](2048)
        scales = 'log:maximal' if name in ('low_freqs', 'high_freqs') else 'log'
        kw = dict(wavelet=('morlet', {'mu': 6.}), scales=scales, l1_norm=False, nv=32)
        print(f"icwt l1_norm=False {name}", flush=True)
        Wx, s = sq.cwt(x.astype(np.float32), t=ts, **kw)
        xrec = sq.icwt(Wx, **kw).cpu().numpy()
        assert mad_rms(x, xrec) < 0.1, (name, mad_rms(x, xrec))
        wav = sq.Wavelet(('morlet', {'mu': 6.}))
        _, _, _, nv = process_scales(scales, len(x), wav, nv=32, get_params=True)
        k = (2 / adm_ssq(wav)) * np.log(2 ** (1 / nv))
        s32 = s.cpu().numpy().astype(np.float32).ravel()
        xo = O.icwt_1int(Wx.cpu().numpy().astype(np.complex128), [(0, len(s32), k, s32**.5)])
        assert np.allclose(xrec, xo, rtol=0, atol=1e-5 * np.abs(xo).max())
