"""trigdiff on the GPU (ssqk_trigdiff): the reference's invariant that it
reproduces cwt's dWx from the padded Wx (z_all_test.py:352-365), the
float64 restatement of utils/common.py:161-245 for every pad mode, and
phase_transform / phase_ssqueeze computing dWx with it."""
import numpy as np
import pytest

from oracle import ssq_oracle as O
from test_gpu_cwt import rel_l2

pytestmark = pytest.mark.gpu


def _trigdiff64(A, padtype, fs=1.):
    """utils/common.py:210-244 in float64 (padding, FFT, i xi fs, inverse
    FFT, unpad to the input's length)."""
    N = A.shape[-1]
    Ap, L, n1, _ = O.pad_signal(A, padtype)
    xi = O.xi_grid(L)
    D = np.fft.ifft(np.fft.fft(Ap, axis=-1) * 1j * xi * fs, axis=-1)
    return D[..., n1:n1 + N]


def test_trigdiff_is_cwt_derivative():
    """z_all_test.py:352-365: trigdiff of the padded Wx equals the dWx of
    cwt(derivative=True) on the unpadded span (here the oracle's float64
    padded transforms; MAE < 1e-7 of the reference's float32 threshold,
    relative)."""
    import torch
    from ssqueezepy_amd.utils import trigdiff
    import ssqueezepy_amd as sq
    from ssqueezepy_amd._ssq_cwt import ssq_cwt_grid
    N = 256
    x = np.random.default_rng(0).standard_normal(N)
    g = ssq_cwt_grid(N, sq.Wavelet('gmw'))
    xp, L, n1, _ = O.pad_signal(x, 'reflect')
    xi = O.xi_grid(L)
    psih = O.psih_fn('gmw')(g['scales64'][:, None] * xi[None])
    psih[:, L // 2] *= 0.5
    X = np.fft.fft(xp)
    Wp = np.fft.ifft(psih * X, axis=-1)
    dWp = np.fft.ifft(psih * X * 1j * xi, axis=-1)
    print("trigdiff vs cwt derivative", flush=True)
    dW2 = trigdiff(torch.as_tensor(Wp.astype(np.complex64), device='cuda'), rpadded=True, N=N,
                   n1=n1).cpu().numpy()
    ref = dWp[:, n1:n1 + N]
    assert dW2.shape == ref.shape
    assert np.mean(np.abs(dW2 - ref)) < 1e-7 * max(1., np.abs(ref).max())
    assert rel_l2(dW2, ref) < 1e-5


@pytest.mark.parametrize('padtype', ['reflect', 'zero', 'symmetric', 'replicate', 'wrap'])
@pytest.mark.parametrize('N', [5, 300, 3000])
def test_trigdiff_vs_float64(padtype, N):
    import torch
    from ssqueezepy_amd.utils import trigdiff
    rng = np.random.default_rng(N)
    A = (rng.standard_normal((6, N)) + 1j * rng.standard_normal((6, N))).astype(np.complex64)
    print(f"trigdiff N={N} {padtype}", flush=True)
    D = trigdiff(torch.as_tensor(A, device='cuda'), fs=2., padtype=padtype, N=N)
    ref = _trigdiff64(A.astype(np.complex128), padtype, fs=2.)
    assert D.shape == ref.shape
    assert rel_l2(D.cpu().numpy(), ref) < 1e-5
    # 3D (batched) rows give the same rows
    D3 = trigdiff(torch.as_tensor(np.stack([A, A[::-1]]), device='cuda'), fs=2.,
                  padtype=padtype, N=N)
    assert torch.equal(D3[0], D) and torch.equal(D3[1], D.flip(0))


def test_phase_ssqueeze_without_dwx():
    """experimental.py:201-209: phase_ssqueeze given only Wx differentiates
    it with trigdiff; the same Tx as passing that dWx."""
    import torch
    import ssqueezepy_amd as sq
    from ssqueezepy_amd.experimental import phase_ssqueeze
    from ssqueezepy_amd.utils import trigdiff
    n = np.arange(2000)
    x = (np.cos(2 * np.pi * (0.02 * n + 0.1 * n**2 / 2000))).astype(np.float32)
    print("phase_ssqueeze without dWx", flush=True)
    Wx, scales = sq.cwt(x, 'gmw')
    s = scales.cpu().numpy()
    Tx1, *_ = phase_ssqueeze(Wx, scales=s, wavelet='gmw', padtype='reflect')
    dWx = trigdiff(Wx, padtype='reflect', N=Wx.shape[-1])
    Tx2, *_ = phase_ssqueeze(Wx, dWx, scales=s, wavelet='gmw')
    assert torch.equal(Tx1, Tx2)
